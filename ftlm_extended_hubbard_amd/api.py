"""Python mirror of the reference's FTLM hot-path API, running on the MI355X through libftlm_hip.so.

Names, argument meaning and error behaviour follow the reference C++ API:
    HubbardParams, nsites, version_string        include/ftlm/hubbard_params.hpp, python/bindings.cpp
    RectLattice, build_hubbard_geometry          include/ftlm/lattice.hpp:25-40
    FockBasis(n_sites, n_up, n_down)             include/ftlm/fock_basis.hpp:11-35
    apply_extended_hubbard(p, basis, hops, pairs, x)   include/ftlm/hubbard_hamiltonian.hpp:13-16
    lanczos_tridiagonal / lanczos_extrema / tridiagonal_extrema   include/ftlm/lanczos.hpp:45-63
    FtlmParams, ftlm_log_partition_complex / _real               include/ftlm/ftlm_thermo.hpp:22-50
Where the reference takes an operator callback (std::function apply_h) this API takes a device
operator (HubbardOperator / KBlockOperator): the operator itself lives on the GPU.
"""

from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


# ---------------------------------------------------------------- params / geometry

@dataclass
class HubbardParams:
    """include/ftlm/hubbard_params.hpp:6-15 (pybind11 defaults of python/bindings.cpp:13-28)."""

    Lx: int = 2
    Ly: int = 2
    t: float = 1.0
    tp: float = 0.0
    U: float = 0.0
    V: float = 0.0
    phi_x: float = 0.0
    phi_y: float = 0.0

    def _c(self) -> L.Params:
        return L.Params(self.Lx, self.Ly, self.t, self.tp, self.U, self.V, self.phi_x, self.phi_y)


def nsites(params: HubbardParams) -> int:
    return L.lib().ftlm_hip_nsites(C.byref(params._c()))


def version_string() -> str:
    return L.lib().ftlm_hip_version().decode()


@dataclass
class RectLattice:
    Lx: int = 1
    Ly: int = 1

    def n_sites(self) -> int:
        return self.Lx * self.Ly

    def site_index(self, x: int, y: int) -> int:
        return y * self.Lx + x

    def coords(self, site: int) -> tuple[int, int]:
        return site % self.Lx, site // self.Lx


def build_hubbard_geometry(p: HubbardParams, lat: RectLattice | None = None):
    """Directed spinful hops and NN pairs, same entries and order as src/lattice.cpp:38-110.
    Returns (hops, pairs) as numpy structured arrays (fields from,to,spin,re,im / i,j)."""
    if lat is not None and (lat.Lx, lat.Ly) != (p.Lx, p.Ly):
        p = HubbardParams(lat.Lx, lat.Ly, p.t, p.tp, p.U, p.V, p.phi_x, p.phi_y)
    n = max(p.Lx * p.Ly, 1)
    cap_h, cap_p = 64 * n + 64, 8 * n + 8
    hops = np.zeros(cap_h, L.HOP_DTYPE)
    pairs = np.zeros(cap_p, L.PAIR_DTYPE)
    nh, np_ = C.c_int(0), C.c_int(0)
    L.check(L.lib().ftlm_hip_build_geometry(C.byref(p._c()), hops.ctypes.data, cap_h, C.byref(nh),
                                            pairs.ctypes.data, cap_p, C.byref(np_)))
    return hops[: nh.value].copy(), pairs[: np_.value].copy()


# ---------------------------------------------------------------- operators

class _Op:
    """Owns one ftlm_op_t handle."""

    _h: C.c_void_p

    def __init__(self):
        self._h = C.c_void_p()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and L._lib is not None:
            L._lib.ftlm_hip_op_destroy(h)
            self._h = C.c_void_p()

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def dim(self) -> int:
        d = C.c_int64(0)
        L.check(L.lib().ftlm_hip_op_dim(self._h, C.byref(d), None, None))
        return d.value

    # ---- H * v
    def apply(self, x: np.ndarray) -> np.ndarray:
        """y = H x for host complex128 vector(s) x of shape (dim,) (copies through the device)."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.shape != (self.dim(),):
            raise ValueError(f"x must have shape ({self.dim()},)")
        y = np.empty_like(x)
        L.check(L.lib().ftlm_hip_apply_host(self._h, x.ctypes.data, y.ctypes.data))
        return y

    def apply_device(self, x, y, R: int = 1, dtype: int = L.FTLM_C128, stream: int | None = None) -> None:
        """y = H x on device buffers (torch tensors or raw pointers), layout v[elem*R + r]."""
        xp = x.data_ptr() if hasattr(x, "data_ptr") else int(x)
        yp = y.data_ptr() if hasattr(y, "data_ptr") else int(y)
        if stream is None:
            try:
                import torch

                stream = torch.cuda.current_stream().cuda_stream
            except Exception:  # noqa: BLE001
                stream = 0
        L.check(L.lib().ftlm_hip_apply(self._h, C.c_void_p(xp), C.c_void_p(yp), R, dtype, C.c_void_p(stream)))

    def apply_pass_times(self, x, y, R: int = 1, dtype: int = L.FTLM_C128, stream: int | None = None):
        """apply_device with the row pass (down bonds) and the column pass (up bonds, diagonal)
        timed by HIP events on the launch stream; synchronises and returns (ms_row, ms_col)."""
        xp = x.data_ptr() if hasattr(x, "data_ptr") else int(x)
        yp = y.data_ptr() if hasattr(y, "data_ptr") else int(y)
        if stream is None:
            import torch

            stream = torch.cuda.current_stream().cuda_stream
        a, b = C.c_double(0.0), C.c_double(0.0)
        L.check(L.lib().ftlm_hip_apply_pass_times(self._h, C.c_void_p(xp), C.c_void_p(yp), R, dtype,
                                                  C.c_void_p(stream), C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- launch shapes
    def autotune(self, R: int, dtype: int = L.FTLM_C64, reps: int = 3, cache: str | None = None) -> dict:
        """Times H*V launch shapes for this operator at (R, dtype) and keeps the fastest
        (ftlm_hip_autotune); every later apply and Lanczos run with that R and dtype uses it.
        With `cache` (a JSON file), a shape stored under this operator's key is restored instead
        of tuned, and a new tuning is added to the file."""
        key = self._shape_key(R, dtype)
        if cache is not None and os.path.exists(cache):
            with open(cache) as f:
                table = json.load(f)
            if key in table:
                self.set_shape(R, dtype, **{k: table[key][k] for k in ("rb", "bpx", "wlog", "grid")})
                return dict(table[key], cached=True)
        sh = L.HvShape()
        stream = 0
        try:
            import torch

            stream = torch.cuda.current_stream().cuda_stream
        except Exception:  # noqa: BLE001
            pass
        L.check(L.lib().ftlm_hip_autotune(self._h, R, dtype, reps, C.c_void_p(stream), C.byref(sh)))
        out = {"rb": sh.rb, "bpx": sh.bpx, "wlog": sh.wlog, "grid": sh.grid, "ms": sh.ms,
               "ms_default": sh.ms_default, "n_tried": sh.n_tried}
        if cache is not None:
            table = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    table = json.load(f)
            table[key] = out
            with open(cache, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
        return dict(out, cached=False)

    def set_shape(self, R: int, dtype: int, rb: int = 0, bpx: int = 0, wlog: int = -1, grid: int = 0) -> None:
        sh = L.HvShape(rb, bpx, wlog, grid, 0.0, 0.0, 0)
        L.check(L.lib().ftlm_hip_set_shape(self._h, R, dtype, C.byref(sh)))

    def get_shape(self, R: int, dtype: int) -> dict:
        sh = L.HvShape()
        L.check(L.lib().ftlm_hip_get_shape(self._h, R, dtype, C.byref(sh)))
        return {"rb": sh.rb, "bpx": sh.bpx, "wlog": sh.wlog, "grid": sh.grid}

    def _shape_key(self, R: int, dtype: int) -> str:
        d, du, dd = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        L.check(L.lib().ftlm_hip_op_dim(self._h, C.byref(d), C.byref(du), C.byref(dd)))
        return f"{type(self).__name__}:{getattr(self, '_tune_tag', '')}:{d.value}:{du.value}:{dd.value}:R{R}:d{dtype}"

    # ---- Lanczos
    def lanczos(self, max_steps: int, seed: int = 1, q0: np.ndarray | None = None):
        """lanczos_tridiagonal (src/lanczos.cpp:125-209): (used, alpha, beta)."""
        m = max(max_steps, 1)
        a, b = np.zeros(m), np.zeros(m)
        nb, used = C.c_int(0), C.c_int(0)
        q = None if q0 is None else np.ascontiguousarray(q0, np.complex128)
        L.check(L.lib().ftlm_hip_lanczos(self._h, max_steps, seed, L.ptr(q), a.ctypes.data, b.ctypes.data,
                                         C.byref(nb), C.byref(used)))
        return used.value, a[: used.value].copy(), b[: nb.value].copy()


class FockBasis:
    """(N_up, N_dn) sector basis, enumerated and ranked on the device (include/ftlm/fock_basis.hpp)."""

    def __init__(self, n_sites: int, n_up: int, n_down: int):
        self._n, self._nu, self._nd = n_sites, n_up, n_down
        self._op = HubbardOperator._create(n_sites, n_up, n_down, None, None, 0.0, 0.0)
        d, du, dd = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        L.check(L.lib().ftlm_hip_op_dim(self._op.handle, C.byref(d), C.byref(du), C.byref(dd)))
        self._dim, self._du, self._dd = d.value, du.value, dd.value
        self._masks = None

    def n_sites(self) -> int:
        return self._n

    def n_up(self) -> int:
        return self._nu

    def n_down(self) -> int:
        return self._nd

    def dim(self) -> int:
        return self._dim

    def masks(self) -> tuple[np.ndarray, np.ndarray]:
        if self._masks is None:
            up = np.zeros(self._dim, np.uint64)
            dn = np.zeros(self._dim, np.uint64)
            L.check(L.lib().ftlm_hip_sector_masks(self._op.handle, 0, self._dim, up.ctypes.data, dn.ctypes.data))
            self._masks = (up, dn)
        return self._masks

    def up_mask(self, idx: int) -> int:
        if not 0 <= idx < self._dim:
            raise IndexError("FockBasis index out of range")
        return int(self.masks()[0][idx])

    def down_mask(self, idx: int) -> int:
        if not 0 <= idx < self._dim:
            raise IndexError("FockBasis index out of range")
        return int(self.masks()[1][idx])

    def index_of(self, up, down):
        """Index of (up, down) or -1; scalars or arrays (ranked on the device)."""
        scalar = np.isscalar(up)
        u = np.atleast_1d(np.asarray(up, dtype=np.uint64))
        d = np.atleast_1d(np.asarray(down, dtype=np.uint64))
        out = np.zeros(len(u), np.int64)
        L.check(L.lib().ftlm_hip_index_of(self._op.handle, u.ctypes.data, d.ctypes.data, len(u), out.ctypes.data))
        return int(out[0]) if scalar else out


class HubbardOperator(_Op):
    """Extended-Hubbard Hamiltonian of one (N_up, N_dn) sector resident on the device."""

    @classmethod
    def _create(cls, n_sites, n_up, n_dn, hops, pairs, U, V):
        self = cls()
        hops = np.zeros(0, L.HOP_DTYPE) if hops is None else np.ascontiguousarray(hops, L.HOP_DTYPE)
        pairs = np.zeros(0, L.PAIR_DTYPE) if pairs is None else np.ascontiguousarray(pairs, L.PAIR_DTYPE)
        L.check(L.lib().ftlm_hip_sector_create(n_sites, n_up, n_dn, hops.ctypes.data if len(hops) else None,
                                               len(hops), pairs.ctypes.data if len(pairs) else None, len(pairs),
                                               float(U), float(V), C.byref(self._h)))
        return self

    @classmethod
    def from_params(cls, p: HubbardParams, n_up: int, n_dn: int) -> "HubbardOperator":
        hops, pairs = build_hubbard_geometry(p)
        op = cls._create(p.Lx * p.Ly, n_up, n_dn, hops, pairs, p.U, p.V)
        op._tune_tag = f"{p.Lx}x{p.Ly}"
        return op


# ---------------------------------------------------------------- translations and k-blocks

def pack_raw_state(up, dn):
    """(up << 16) | dn (include/ftlm/symmetry/packed_raw_state.hpp:12-14), elementwise."""
    return ((np.asarray(up, np.uint32) << np.uint32(16)) | np.asarray(dn, np.uint32)).astype(np.uint32)


def translate_states(Lx: int, Ly: int, states, dx: int, dy: int):
    """translate_occupation_bits_rect on both strings and fermionic_translation_sign_rect
    (occup_bits_translate.hpp:10-25, fermionic_translation.hpp:32-54): (translated, sign)."""
    st = np.ascontiguousarray(states, np.uint32)
    out, sg = np.zeros(len(st), np.uint32), np.zeros(len(st))
    L.check(L.lib().ftlm_hip_translate_states(Lx, Ly, st.ctypes.data, len(st), dx, dy, out.ctypes.data,
                                              sg.ctypes.data))
    return out, sg


def canonical_states(Lx: int, Ly: int, states):
    """canonical_raw_state_rect and |translation_stabilizer_rect| (orbit.hpp:93-155)."""
    st = np.ascontiguousarray(states, np.uint32)
    out, ss = np.zeros(len(st), np.uint32), np.zeros(len(st), np.int32)
    L.check(L.lib().ftlm_hip_canonical_states(Lx, Ly, st.ctypes.data, len(st), out.ctypes.data, ss.ctypes.data))
    return out, ss


@dataclass(frozen=True)
class MomentumSector:
    """MomentumSector (include/ftlm/symmetry/momentum_sector.hpp:20-33)."""
    kx: int = 0
    ky: int = 0
    lx: int = 4
    ly: int = 4

    def flat_index(self) -> int:
        return self.ky * self.lx + self.kx


def translation_bloch_phase(K: MomentumSector, dx: int, dy: int) -> complex:
    """exp(-2 pi i (kx dx / lx + ky dy / ly)) (momentum_sector.hpp:57-63)."""
    re, im = C.c_double(0.0), C.c_double(0.0)
    L.check(L.lib().ftlm_hip_bloch_phase(K.lx, K.ly, K.kx, K.ky, dx, dy, C.byref(re), C.byref(im)))
    return complex(re.value, im.value)


@dataclass
class OrbitKRepEntries:
    rep: np.ndarray          # packed representatives, sorted
    orbit_size: np.ndarray
    stabilizer_size: np.ndarray
    A_k: np.ndarray          # complex


class KBlockOperator(_Op):
    """OrbitKBlockBasis (src/orbit_k_block_basis.cpp:70-149) with the operator of
    apply_orbit_k_block (:151-220), resident on the device.  Blocks of one sector and
    Hamiltonian share one orbit graph when created with `share=`."""

    @classmethod
    def _create(cls, Lx, Ly, n_up, n_dn, K: MomentumSector, hops, pairs, U, V, tol=1e-12, share=None):
        self = cls()
        hops = np.ascontiguousarray(hops, L.HOP_DTYPE)
        pairs = np.ascontiguousarray(pairs, L.PAIR_DTYPE)
        L.check(L.lib().ftlm_hip_kblock_create(Lx, Ly, n_up, n_dn, K.kx, K.ky, float(tol),
                                               hops.ctypes.data if len(hops) else None, len(hops),
                                               pairs.ctypes.data if len(pairs) else None, len(pairs), float(U),
                                               float(V), None if share is None else share._h, C.byref(self._h)))
        self.K = K
        self._share = share  # keeps the shared graph's owner alive for readability; the graph is refcounted
        return self

    @classmethod
    def from_params(cls, p: HubbardParams, n_up: int, n_dn: int, K: MomentumSector, tol: float = 1e-12,
                    share: "KBlockOperator | None" = None) -> "KBlockOperator":
        hops, pairs = build_hubbard_geometry(p)
        K = MomentumSector(K.kx, K.ky, p.Lx, p.Ly)
        op = cls._create(p.Lx, p.Ly, n_up, n_dn, K, hops, pairs, p.U, p.V, tol, share)
        op._tune_tag = f"{p.Lx}x{p.Ly}:{n_up},{n_dn}:K{K.kx},{K.ky}"
        return op

    def info(self):
        kx, ky, no, nnz = C.c_int(0), C.c_int(0), C.c_int64(0), C.c_int64(0)
        L.check(L.lib().ftlm_hip_kblock_info(self._h, C.byref(kx), C.byref(ky), C.byref(no), C.byref(nnz)))
        return {"kx": kx.value, "ky": ky.value, "n_orbits": no.value, "nnz": nnz.value}

    def entries(self) -> OrbitKRepEntries:
        n = self.dim()
        rep, osz, ssz = np.zeros(n, np.uint32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        A = np.zeros(2 * n)
        if n:
            L.check(L.lib().ftlm_hip_kblock_entries(self._h, 0, n, rep.ctypes.data, osz.ctypes.data, ssz.ctypes.data,
                                                    A.ctypes.data))
        return OrbitKRepEntries(rep, osz, ssz, A[0::2] + 1j * A[1::2])

    def reduced_index_of_rep(self, reps) -> np.ndarray:
        r = np.ascontiguousarray(np.atleast_1d(reps), np.uint32)
        out = np.zeros(len(r), np.int64)
        L.check(L.lib().ftlm_hip_kblock_index_of_rep(self._h, r.ctypes.data, len(r), out.ctypes.data))
        return out

    def map_raw_to_rep(self, states):
        """(index, dx, dy, eta) per packed state; index -1 where the reference returns false."""
        st = np.ascontiguousarray(np.atleast_1d(states), np.uint32)
        n = len(st)
        idx, dx, dy, eta = np.zeros(n, np.int64), np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        L.check(L.lib().ftlm_hip_kblock_map(self._h, st.ctypes.data, n, idx.ctypes.data, dx.ctypes.data,
                                            dy.ctypes.data, eta.ctypes.data))
        return idx, dx, dy, eta


class KStackOperator(_Op):
    """The momentum blocks of one sector and Hamiltonian side by side (ftlm_hip_kstack_create):
    the reference's per-K loop over apply_orbit_k_block (benchmarks/bench_ftlm_nmu_4x4_k.cpp:801-1103)
    as one operator over the orbits.  Its applies take R = len(blocks) * rk vectors laid out
    [orbit][block][rk]; `scatter`/`gather` move one block's [dim_j][rk] vectors in and out."""

    @classmethod
    def from_blocks(cls, blocks) -> "KStackOperator":
        self = cls()
        arr = (C.c_void_p * len(blocks))(*[b.handle.value for b in blocks])
        L.check(L.lib().ftlm_hip_kstack_create(arr, len(blocks), C.byref(self._h)))
        self._blocks = list(blocks)  # the stack borrows the blocks' row maps
        return self

    @property
    def n_blocks(self) -> int:
        return len(self._blocks)

    def _copy(self, fn, j, src, dst, rk, dtype, stream):
        sp = src.data_ptr() if hasattr(src, "data_ptr") else int(src)
        dp = dst.data_ptr() if hasattr(dst, "data_ptr") else int(dst)
        if stream is None:
            import torch

            stream = torch.cuda.current_stream().cuda_stream
        L.check(fn(self._h, j, C.c_void_p(sp), rk, dtype, C.c_void_p(dp), C.c_void_p(stream)))

    def scatter(self, j: int, xb, xs, rk: int = 1, dtype: int = L.FTLM_C128, stream: int | None = None) -> None:
        """Block j's vectors xb [dim_j][rk] into slots (j, 0..rk) of the stacked block xs."""
        self._copy(L.lib().ftlm_hip_kstack_scatter, j, xb, xs, rk, dtype, stream)

    def gather(self, j: int, xs, xb, rk: int = 1, dtype: int = L.FTLM_C128, stream: int | None = None) -> None:
        """Slots (j, 0..rk) of the stacked block xs back into block j's layout xb [dim_j][rk]."""
        self._copy(L.lib().ftlm_hip_kstack_gather, j, xs, xb, rk, dtype, stream)


_APPLY_CACHE: dict = {}


class KGramOperator(_Op):
    """HubbardMomentumBlock (src/hubbard_momentum_action.cpp:37-89) in the Gram basis of
    MomentumPhiGramBasis (src/orbit_bloch_phi.cpp:226-355): y = Phi^dag H Phi x through the full
    sector.  `sector` (a HubbardOperator of the same lattice and particle numbers) may be shared by
    the blocks of every K; it is kept alive by the block."""

    @classmethod
    def from_params(cls, p: HubbardParams, n_up: int, n_dn: int, K: MomentumSector,
                    sector: "HubbardOperator | None" = None,
                    share: "KBlockOperator | None" = None) -> "KGramOperator":
        sector = HubbardOperator.from_params(p, n_up, n_dn) if sector is None else sector
        kb = KBlockOperator.from_params(p, n_up, n_dn, K, share=share)
        self = cls()
        L.check(L.lib().ftlm_hip_kgram_create(sector._h, kb._h, C.byref(self._h)))
        self.sector, self.K = sector, MomentumSector(K.kx, K.ky, p.Lx, p.Ly)
        return self


def apply_extended_hubbard(p: HubbardParams, basis: FockBasis, hops, pairs, x: np.ndarray) -> np.ndarray:
    """Drop-in of apply_extended_hubbard (src/hubbard_hamiltonian.cpp:51-108): returns y = H x.
    The device operator for (basis, hops, pairs, U, V) is built once and cached."""
    hops = np.ascontiguousarray(hops, L.HOP_DTYPE)
    pairs = np.ascontiguousarray(pairs, L.PAIR_DTYPE)
    key = (basis.n_sites(), basis.n_up(), basis.n_down(), hops.tobytes(), pairs.tobytes(), float(p.U), float(p.V))
    op = _APPLY_CACHE.get(key)
    if op is None:
        if len(_APPLY_CACHE) > 16:
            _APPLY_CACHE.clear()
        op = HubbardOperator._create(basis.n_sites(), basis.n_up(), basis.n_down(), hops, pairs, p.U, p.V)
        _APPLY_CACHE[key] = op
    return op.apply(x)


# ---------------------------------------------------------------- Lanczos / FTLM

@dataclass
class LanczosExtrema:
    min_eval: float = 0.0
    max_eval: float = 0.0
    steps_used: int = 0


def lanczos_tridiagonal(op: _Op, max_steps: int, seed: int, q0: np.ndarray | None = None):
    """(used, alpha, beta) as src/lanczos.cpp:125-209; q0 is the optional supplied start vector."""
    return op.lanczos(max_steps, seed, q0)


def lanczos_extrema(op: _Op, max_steps: int, seed: int = 1) -> LanczosExtrema:
    lo, hi, used = C.c_double(0), C.c_double(0), C.c_int(0)
    L.check(L.lib().ftlm_hip_lanczos_extrema(op.handle, max_steps, seed, C.byref(lo), C.byref(hi), C.byref(used)))
    return LanczosExtrema(lo.value, hi.value, used.value)


def tridiagonal_extrema(alpha, beta, steps: int) -> tuple[float, float]:
    a = np.ascontiguousarray(alpha[:steps], np.float64)
    b = np.ascontiguousarray(beta, np.float64) if len(beta) else np.zeros(1)
    lo, hi = C.c_double(0), C.c_double(0)
    L.check(L.lib().ftlm_hip_tridiagonal_extrema(a.ctypes.data if steps > 0 else None, steps, b.ctypes.data,
                                                 len(beta), C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


@dataclass
class FtlmParams:
    """include/ftlm/ftlm_thermo.hpp:22-32, plus the device-path options."""

    n_random: int = 16
    lanczos_steps: int = 96
    seed: int = 1
    start: int = L.FTLM_START_GAUSS_MT
    dtype: int = L.FTLM_C128
    quad: int = L.FTLM_QUAD_REFERENCE
    clamp_M_to_dim: bool = False
    path: int = L.FTLM_PATH_AUTO
    reorth: int = L.FTLM_REORTH_NONE  # FTLM_REORTH_*: the reference never reorthogonalises

    def _c(self, real: bool = False) -> L.FtlmOpts:
        return L.FtlmOpts(self.n_random, self.lanczos_steps, self.seed, self.start, self.dtype, self.quad,
                          int(self.clamp_M_to_dim), int(real), int(self.path), int(self.reorth))


def _log_partition(op: _Op, betas, par: FtlmParams, real: bool):
    scalar = np.isscalar(betas)
    b = np.atleast_1d(np.asarray(betas, np.float64))
    out = np.zeros(len(b))
    opts = par._c(real)
    L.check(L.lib().ftlm_hip_ftlm_log_partition(op.handle, b.ctypes.data, len(b), C.byref(opts), out.ctypes.data))
    return float(out[0]) if scalar else out


def ftlm_log_partition_complex(op: _Op, beta, par: FtlmParams = FtlmParams()):
    """src/ftlm_thermo.cpp:213-256; `beta` may be an array (the Lanczos data are reused)."""
    return _log_partition(op, beta, par, False)


def ftlm_log_partition_real(op: _Op, beta, par: FtlmParams = FtlmParams()):
    """src/ftlm_thermo.cpp:150-211 with the Re(H) apply of benchmarks/bench_ftlm_nmu_rect.cpp:151-159."""
    return _log_partition(op, beta, par, True)


def exact_log_partition(op: _Op, betas, return_evals: bool = False):
    """exact_log_partition_small_complex (benchmarks/bench_ftlm_nmu_4x4_k.cpp:249-336) on the
    device for an operator of at most 1024 states: lnZ per beta (and the 2*dim Jacobi diagonal of
    the real embedding when return_evals)."""
    scalar = np.isscalar(betas)
    b = np.atleast_1d(np.asarray(betas, np.float64))
    out = np.zeros(len(b))
    ev = np.zeros(max(2 * op.dim(), 2)) if return_evals else None
    L.check(L.lib().ftlm_hip_exact_log_partition(op.handle, b.ctypes.data, len(b), out.ctypes.data, L.ptr(ev)))
    res = float(out[0]) if scalar else out
    return (res, ev[: 2 * op.dim()]) if return_evals else res


# ---------------------------------------------------------------- derivatives

GRAD_PARAMS = ("t", "tp", "U", "V", "phi_x", "phi_y")


def geometry_derivative(p: HubbardParams, which: str):
    """d(coefficient)/d(theta) of every hop of build_hubbard_geometry, theta in t, tp, phi_x, phi_y."""
    code = {"t": 0, "tp": 1, "phi_x": 2, "phi_y": 3}[which]
    n = max(p.Lx * p.Ly, 1)
    cap = 64 * n + 64
    hops = np.zeros(cap, L.HOP_DTYPE)
    nh = C.c_int(0)
    L.check(L.lib().ftlm_hip_build_geometry_derivative(C.byref(p._c()), code, hops.ctypes.data, cap, C.byref(nh)))
    return hops[: nh.value].copy()


def derivative_operator(p: HubbardParams, n_up: int, n_dn: int, which: str,
                        K: "MomentumSector | None" = None) -> "_Op":
    """dH/dtheta on the (n_up, n_dn) sector, or on its momentum-K block, as an operator: derivative
    hops for t, tp, phi_x, phi_y; the double-occupancy (U) or nearest-neighbour density (V) term for
    U, V.  dH commutes with the lattice translations, so its K block is the block of dH."""
    n = p.Lx * p.Ly
    no_hops = np.zeros(0, L.HOP_DTYPE)
    if which == "U":
        hops, pairs, U, V = no_hops, None, 1.0, 0.0
    elif which == "V":
        hops, (_, pairs), U, V = no_hops, build_hubbard_geometry(p), 0.0, 1.0
    else:
        hops, pairs, U, V = geometry_derivative(p, which), None, 0.0, 0.0
    if K is None:
        return HubbardOperator._create(n, n_up, n_dn, hops if len(hops) else None, pairs, U, V)
    K = MomentumSector(K.kx, K.ky, p.Lx, p.Ly)
    pairs = np.zeros(0, L.PAIR_DTYPE) if pairs is None else pairs
    return KBlockOperator._create(p.Lx, p.Ly, n_up, n_dn, K, hops, pairs, U, V)


def ftlm_gradient(op: _Op, dops, betas, par: FtlmParams):
    """lnZ(beta) of the FTLM estimator with the converged Gauss rule and its exact derivatives
    along the operators dops (dH/dtheta): returns (lnZ [nb], grad [len(dops)][nb])."""
    b = np.ascontiguousarray(np.atleast_1d(betas), np.float64)
    nb, npar = len(b), len(dops)
    lnz, g = np.zeros(nb), np.zeros((max(npar, 1), nb))
    arr = (C.c_void_p * max(npar, 1))(*[d.handle.value for d in dops]) if npar else None
    L.check(L.lib().ftlm_hip_ftlm_gradient(op.handle, arr, npar, C.byref(par._c()), b.ctypes.data, nb,
                                           lnz.ctypes.data, g.ctypes.data))
    return lnz, g[:npar]


def ftlm_vjp(op: _Op, dops, betas, weights, par: FtlmParams, checkpoint: int = 0, mode: int = L.FTLM_TS_AUTO):
    """Reverse mode: (lnZ [nb], g [len(dops)]) with g_p = sum_b weights[b] dlnZ(beta_b)/dtheta_p,
    dH/dtheta_p given by the operators dops; Krylov vectors are recomputed from checkpoints every
    `checkpoint` steps (0: round(sqrt(2M)))."""
    b = np.ascontiguousarray(np.atleast_1d(betas), np.float64)
    gw = np.ascontiguousarray(np.broadcast_to(np.asarray(weights, np.float64), b.shape))
    nb, npar = len(b), len(dops)
    lnz, g = np.zeros(nb), np.zeros(max(npar, 1))
    arr = (C.c_void_p * max(npar, 1))(*[d.handle.value for d in dops]) if npar else None
    L.check(L.lib().ftlm_hip_ftlm_vjp(op.handle, arr, npar, C.byref(par._c()), b.ctypes.data, nb, gw.ctypes.data,
                                      int(checkpoint), int(mode), lnz.ctypes.data, g.ctypes.data))
    return lnz, g[:npar]


def ftlm_observables(op: _Op, aops, betas, par: FtlmParams):
    """FTLM estimates <A_p>(beta) over the samples of ftlm_gradient (same seeds and steps, converged
    Gauss rule): (lnZ [nb], obs [len(aops)][nb])."""
    b = np.ascontiguousarray(np.atleast_1d(betas), np.float64)
    nb, n = len(b), len(aops)
    lnz, obs = np.zeros(nb), np.zeros((n, nb))
    arr = (C.c_void_p * n)(*[a.handle.value for a in aops])
    L.check(L.lib().ftlm_hip_ftlm_observables(op.handle, arr, n, C.byref(par._c()), b.ctypes.data, nb,
                                              lnz.ctypes.data, obs.ctypes.data))
    return lnz, obs


def ftlm_log_partition_observables(p: HubbardParams, n_up: int, n_dn: int, betas, par: FtlmParams,
                                   params=("t", "tp", "U", "V"), K: "MomentumSector | None" = None):
    """lnZ(beta) and the FTLM estimates of <dH/dtheta> for theta in params (t, tp, U, V, phi_x,
    phi_y): <D> for U, the NN pair count for V, the hop parts for t and t'.  At exact
    thermodynamics these are -(1/beta) dlnZ/dtheta."""
    op = HubbardOperator.from_params(p, n_up, n_dn) if K is None else KBlockOperator.from_params(p, n_up, n_dn, K)
    aops = [derivative_operator(p, n_up, n_dn, q, K) for q in params]
    lnz, obs = ftlm_observables(op, aops, betas, par)
    return lnz, {q: obs[i] for i, q in enumerate(params)}


class FtlmTape:
    """The forward pass of the adjoint method (ftlm_hip_ftlm_tape_forward): lnZ(betas) now, then
    backward(weights) -> g_p = sum_b weights[b] dlnZ(beta_b)/dtheta_p as often as needed.  Holds
    the operators and the device checkpoints until closed."""

    def __init__(self, op: _Op, dops, betas, par: FtlmParams, checkpoint: int = 0, mode: int = L.FTLM_TS_AUTO):
        self.betas = np.ascontiguousarray(np.atleast_1d(betas), np.float64)
        self._ops = (op, list(dops))
        self._h = C.c_void_p()
        self.lnZ = np.zeros(len(self.betas))
        npar = len(dops)
        arr = (C.c_void_p * max(npar, 1))(*[d.handle.value for d in dops]) if npar else None
        L.check(L.lib().ftlm_hip_ftlm_tape_forward(op.handle, arr, npar, C.byref(par._c()), self.betas.ctypes.data,
                                                   len(self.betas), int(checkpoint), int(mode),
                                                   self.lnZ.ctypes.data, C.byref(self._h)))

    def backward(self, weights) -> np.ndarray:
        if not self._h:
            raise ValueError("the tape is closed")
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(weights, np.float64), self.betas.shape))
        g = np.zeros(max(len(self._ops[1]), 1))
        L.check(L.lib().ftlm_hip_ftlm_tape_backward(self._h, w.ctypes.data, g.ctypes.data))
        return g[: len(self._ops[1])]

    def close(self) -> None:
        if self._h:
            L.lib().ftlm_hip_ftlm_tape_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ftlm_log_partition_vjp(p: HubbardParams, n_up: int, n_dn: int, betas, weights, par: FtlmParams,
                           params=("t", "tp", "U", "V", "phi_x"), mu: float | None = None,
                           K: "MomentumSector | None" = None, checkpoint: int = 0, mode: int = L.FTLM_TS_AUTO):
    """lnZ(beta) of the sector (or of its momentum-K block) and sum_b weights[b] dlnZ(beta_b)/dtheta
    for theta in params (any of t, tp, U, V, phi_x, phi_y, mu), by the adjoint Lanczos sweep."""
    op = HubbardOperator.from_params(p, n_up, n_dn) if K is None else KBlockOperator.from_params(p, n_up, n_dn, K)
    hop = [q for q in params if q != "mu"]
    dops = [derivative_operator(p, n_up, n_dn, q, K) for q in hop]
    b = np.atleast_1d(np.asarray(betas, np.float64))
    w = np.broadcast_to(np.asarray(weights, np.float64), b.shape)
    lnz, g = ftlm_vjp(op, dops, b, w, par, checkpoint, mode)
    out = {q: float(g[i]) for i, q in enumerate(hop)}
    if mu is not None:
        lnz = lnz + b * mu * (n_up + n_dn)
    if "mu" in params:
        out["mu"] = float(np.sum(w * b) * (n_up + n_dn))
    return lnz, out


def ftlm_log_partition_gradient(p: HubbardParams, n_up: int, n_dn: int, betas, par: FtlmParams,
                                params=("t", "tp", "U", "V", "phi_x"), mu: float | None = None,
                                K: "MomentumSector | None" = None):
    """lnZ(beta) of the sector (or of its momentum-K block) and d lnZ / d theta for theta in params
    (any of t, tp, U, V, phi_x, phi_y, mu).  With mu, H - mu N is used: lnZ gains beta mu N and
    d lnZ / d mu = beta N."""
    op = HubbardOperator.from_params(p, n_up, n_dn) if K is None else KBlockOperator.from_params(p, n_up, n_dn, K)
    hop = [q for q in params if q != "mu"]
    dops = [derivative_operator(p, n_up, n_dn, q, K) for q in hop]
    b = np.atleast_1d(np.asarray(betas, np.float64))
    lnz, g = ftlm_gradient(op, dops, b, par)
    out = {q: g[i] for i, q in enumerate(hop)}
    if mu is not None:
        lnz = lnz + b * mu * (n_up + n_dn)
    if "mu" in params:
        out["mu"] = b * (n_up + n_dn)
    return lnz, out


@dataclass
class RitzData:
    ritz: np.ndarray     # [n_ops, R, M]
    weight: np.ndarray   # [n_ops, R, M]
    used: np.ndarray     # [n_ops, R]
    dims: np.ndarray     # [n_ops]
    alpha: np.ndarray | None = None
    beta: np.ndarray | None = None


def ftlm_ritz(ops, par: FtlmParams, real: bool = False, keep_tridiagonals: bool = False) -> RitzData:
    n = len(ops)
    R, M = max(1, par.n_random), max(1, par.lanczos_steps)
    rz, wt = np.zeros((n, R, M)), np.zeros((n, R, M))
    us = np.zeros((n, R), np.int32)
    al = np.zeros((n, R, M)) if keep_tridiagonals else None
    be = np.zeros((n, R, M)) if keep_tridiagonals else None
    arr = (C.c_void_p * n)(*[o.handle.value for o in ops])
    opts = par._c(real)
    L.check(L.lib().ftlm_hip_ftlm_ritz(arr, n, C.byref(opts), rz.ctypes.data, wt.ctypes.data, us.ctypes.data,
                                       L.ptr(al), L.ptr(be), None))
    dims = np.array([o.dim() for o in ops], np.int64)
    return RitzData(rz, wt, us, dims, al, be)


def ftlm_thermo(rd: RitzData, betas, N=None, mus=None):
    """Per-operator lnZ(beta), E(beta) and, with N and mus, the grand-canonical lnZ, <N>, E."""
    n, R, M = rd.ritz.shape
    b = np.ascontiguousarray(np.atleast_1d(betas), np.float64)
    nb = len(b)
    lnz, e = np.zeros((n, nb)), np.zeros((n, nb))
    nmu = 0 if mus is None else len(mus)
    mu = None if mus is None else np.ascontiguousarray(mus, np.float64)
    Nn = None if N is None else np.ascontiguousarray(N, np.int32)
    lg, ng, eg = (np.zeros((nb, max(nmu, 1))) for _ in range(3))
    L.check(L.lib().ftlm_hip_ftlm_thermo(rd.ritz.ctypes.data, rd.weight.ctypes.data, rd.used.ctypes.data, n, R, M,
                                         rd.dims.ctypes.data, L.ptr(Nn), b.ctypes.data, nb, L.ptr(mu), nmu,
                                         lnz.ctypes.data, e.ctypes.data, lg.ctypes.data, ng.ctypes.data,
                                         eg.ctypes.data))
    out = {"lnZ": lnz, "E": e}
    if nmu:
        out.update({"lnZ_gc": lg, "N_gc": ng, "E_gc": eg})
    return out


class KrylovSession:
    """R device-resident Lanczos recurrences advanced step by step (LanczosComplexWorkspace,
    include/ftlm/lanczos.hpp:13-35, holding R samples)."""

    def __init__(self, op: _Op, R: int, max_steps: int, dtype: int = L.FTLM_C128,
                 start: int = L.FTLM_START_PHASE, seed: int = 1):
        self._op = op  # keeps the operator alive
        self.R, self.M = R, max_steps
        self._h = C.c_void_p()
        L.check(L.lib().ftlm_hip_krylov_create(op.handle, R, max_steps, dtype, start, seed, C.byref(self._h)))

    def run(self, n_steps: int, stream: int | None = None, graph: bool = False) -> None:
        """Advance n_steps on `stream`; graph=True captures them as one hipGraph first."""
        if stream is None:
            try:
                import torch

                stream = torch.cuda.current_stream().cuda_stream
            except Exception:  # noqa: BLE001
                stream = 0
        fn = L.lib().ftlm_hip_krylov_run_graph if graph else L.lib().ftlm_hip_krylov_run
        L.check(fn(self._h, n_steps, C.c_void_p(stream)))

    def result(self):
        a, b = np.zeros((self.M, self.R)), np.zeros((self.M, self.R))
        used, alive = np.zeros(self.R, np.int32), np.zeros(self.R, np.int32)
        L.check(L.lib().ftlm_hip_krylov_result(self._h, a.ctypes.data, b.ctypes.data, used.ctypes.data,
                                               alive.ctypes.data))
        return a, b, used, alive

    def nbytes(self) -> int:
        v = C.c_int64(0)
        L.check(L.lib().ftlm_hip_krylov_bytes(self._h, C.byref(v)))
        return v.value

    def close(self) -> None:
        if self._h is not None and self._h.value and L._lib is not None:
            L._lib.ftlm_hip_krylov_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        self.close()


def quadrature_ritz(alpha, beta, quad: int = L.FTLM_QUAD_REFERENCE):
    a = np.ascontiguousarray(alpha, np.float64)
    b = np.ascontiguousarray(beta if len(beta) else np.zeros(1), np.float64)
    n = len(a)
    lam, w = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    L.check(L.lib().ftlm_hip_quadrature_ritz(a.ctypes.data, n, b.ctypes.data, len(beta), quad, lam.ctypes.data,
                                             w.ctypes.data))
    return lam[:n], w[:n]


def device_count() -> int:
    c = C.c_int(0)
    L.check(L.lib().ftlm_hip_device_count(C.byref(c)))
    return c.value


__all__ = [
    "HubbardParams", "nsites", "version_string", "RectLattice", "build_hubbard_geometry", "FockBasis",
    "HubbardOperator", "apply_extended_hubbard", "LanczosExtrema", "lanczos_tridiagonal", "lanczos_extrema",
    "tridiagonal_extrema", "FtlmParams", "ftlm_log_partition_complex", "ftlm_log_partition_real", "RitzData",
    "ftlm_ritz", "ftlm_thermo", "quadrature_ritz", "device_count", "KrylovSession",
]
