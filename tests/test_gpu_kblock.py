"""GPU parity of the translation primitives and the orbit k-blocks (through the C ABI) against the
reference's golden vectors, the CPU oracle and identities the symmetry reduction must satisfy.

Tolerances: packed states, orbit/stabiliser sizes, row indices, translations and signs bit-exact;
A_k within 1e-14; complex128 H_K x within 1e-12 relative; block spectra within 1e-9 of the full
sector's; lnZ within 1e-9 absolute.
"""

import sys

import numpy as np
import pytest

from pyoracle import params8

pytestmark = pytest.mark.gpu

api = pytest.importorskip("ftlm_extended_hubbard_amd.api")
from ftlm_extended_hubbard_amd import _lib as L  # noqa: E402


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _device():
    if api.device_count() < 1:
        pytest.skip("no HIP device")


def model(lx, ly, phi=0.0):
    # the orbit fixtures' Hamiltonian (tests/golden/make_golden.py)
    return api.HubbardParams(lx, ly, 1.0, -0.25, 8.0, 1.0, phi, 0.0)


def kblock(lx, ly, nu, nd, kx, ky, share=None, phi=0.0):
    return api.KBlockOperator.from_params(model(lx, ly, phi), nu, nd, api.MomentumSector(kx, ky, lx, ly),
                                          share=share)


def sector_states(oracle, lx, ly, nu, nd):
    up, dn = oracle.basis_masks(lx * ly, nu, nd)
    return api.pack_raw_state(up, dn)


def test_translate_and_canonical_vs_oracle(oracle):
    rng = np.random.default_rng(5)
    for (lx, ly, nu, nd) in [(4, 4, 8, 8), (4, 2, 3, 2), (3, 2, 3, 3), (2, 2, 2, 1), (4, 3, 6, 5)]:
        say("translate", lx, ly, nu, nd)
        st = sector_states(oracle, lx, ly, nu, nd)
        st = st[rng.choice(len(st), size=min(len(st), 300), replace=False)]
        for dx in range(lx):
            for dy in range(ly):
                out, sg = api.translate_states(lx, ly, st, dx, dy)
                want = [oracle.translate(int(s) >> 16, int(s) & 0xFFFF, lx, ly, dx, dy) for s in st]
                wsg = [oracle.translation_sign(int(s) >> 16, int(s) & 0xFFFF, lx, ly, dx, dy) for s in st]
                assert np.array_equal(out, np.array(want, np.uint32)), (lx, ly, dx, dy)
                assert np.array_equal(sg, np.array(wsg)), (lx, ly, dx, dy)
        can, ss = api.canonical_states(lx, ly, st)
        want = [oracle.canonical(int(s) >> 16, int(s) & 0xFFFF, lx, ly) for s in st]
        assert np.array_equal(can, np.array(want, np.uint32))
        # |orbit| * |stabiliser| = |G|: the stabiliser counts the translations fixing the state
        fixed = np.zeros(len(st), np.int32)
        for dx in range(lx):
            for dy in range(ly):
                fixed += (api.translate_states(lx, ly, st, dx, dy)[0] == st)
        assert np.array_equal(ss, fixed)


def test_bloch_phase():
    K = api.MomentumSector(1, 2, 4, 4)
    assert abs(api.translation_bloch_phase(K, 0, 0) - 1.0) == 0.0
    z = api.translation_bloch_phase(K, 1, 1)
    assert abs(z - np.exp(-2j * np.pi * (1 / 4 + 2 / 4))) < 1e-15


@pytest.mark.parametrize("lx,ly,nu,nd", [(2, 2, 2, 2), (3, 2, 3, 3), (4, 2, 4, 4), (4, 2, 2, 3)])
def test_kblock_basis_golden(golden, lx, ly, nu, nd):
    say("kblock basis", lx, ly, nu, nd)
    dims = golden[f"orbit/{lx}x{ly}/{nu}_{nd}/dims"]
    first = None
    for ky in range(ly):
        for kx in range(lx):
            B = kblock(lx, ly, nu, nd, kx, ky, share=first)
            first = first or B
            assert B.dim() == dims[ky, kx], (kx, ky)
            key = f"orbit/{lx}x{ly}/{nu}_{nd}/k{kx}{ky}"
            if key + "/reps" not in golden.files:
                continue
            e = B.entries()
            assert np.array_equal(e.rep, golden[key + "/reps"])
            assert np.array_equal(e.orbit_size, golden[key + "/orbit_size"])
            assert np.array_equal(e.stabilizer_size, golden[key + "/stab_size"])
            assert np.max(np.abs(e.A_k - golden[key + "/A"])) < 1e-14
            assert np.array_equal(B.reduced_index_of_rep(e.rep), np.arange(B.dim()))


@pytest.mark.parametrize("lx,ly,nu,nd,kx,ky", [(3, 2, 3, 3, 1, 0), (2, 2, 2, 2, 0, 0), (4, 2, 4, 4, 2, 1),
                                               (4, 2, 2, 3, 1, 0)])
def test_kblock_map_apply_lnz_golden(golden, oracle, lx, ly, nu, nd, kx, ky):
    say("kblock apply", lx, ly, nu, nd, kx, ky)
    key = f"orbit_apply/{lx}x{ly}/{nu}_{nd}/k{kx}{ky}"
    B = kblock(lx, ly, nu, nd, kx, ky)
    st = sector_states(oracle, lx, ly, nu, nd)
    idx, dx, dy, eta = B.map_raw_to_rep(st)
    want = golden[key + "/map"]
    assert np.array_equal(idx, want[:, 0])
    ok = want[:, 0] >= 0
    assert np.array_equal(dx[ok], want[ok, 1]) and np.array_equal(dy[ok], want[ok, 2])
    assert np.array_equal(eta[ok].astype(np.int32), want[ok, 3])
    y = B.apply(golden[key + "/x"])
    assert rel(y, golden[key + "/y"]) < 1e-12
    par = api.FtlmParams(n_random=2, lanczos_steps=12, seed=3)
    lnz = api.ftlm_log_partition_complex(B, 1.5, par)
    w = float(golden[key + "/lnZ_b1.5_R2_M12_s3"][0])
    assert abs(lnz - w) < 1e-9 * max(1.0, abs(w)), (lnz, w)


def test_kblock_apply_vs_oracle_flux(oracle):
    """Nonzero flux breaks translation symmetry; the block operator is still the reference's
    push-form matrix (not Hermitian), which the pull table must reproduce."""
    lx, ly, nu, nd, kx, ky = 4, 2, 2, 3, 1, 1
    say("kblock flux", lx, ly, nu, nd, kx, ky)
    B = kblock(lx, ly, nu, nd, kx, ky, phi=0.3)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(B.dim()) + 1j * rng.standard_normal(B.dim())
    y = B.apply(x)
    want = oracle.orbit_apply(params8(lx, ly, 1.0, -0.25, 8.0, 1.0, 0.3, 0.0), nu, nd, kx, ky, x)
    assert rel(y, want) < 1e-12


def test_kblock_block_vectors_and_c64(oracle):
    import torch

    lx, ly, nu, nd = 4, 2, 4, 4
    say("kblock R/c64", lx, ly, nu, nd)
    B = kblock(lx, ly, nu, nd, 1, 0)
    d = B.dim()
    rng = np.random.default_rng(2)
    for R in (1, 3, 8, 16):
        X = rng.standard_normal((d, R)) + 1j * rng.standard_normal((d, R))
        xd = torch.tensor(X, dtype=torch.complex128, device="cuda").contiguous()
        yd = torch.empty_like(xd)
        B.apply_device(xd, yd, R=R, dtype=L.FTLM_C128)
        Y = yd.cpu().numpy()
        for r in range(R):
            assert np.array_equal(Y[:, r], B.apply(X[:, r])), (R, r)
        x64 = xd.to(torch.complex64)
        y64 = torch.empty_like(x64)
        B.apply_device(x64, y64, R=R, dtype=L.FTLM_C64)
        assert rel(y64.cpu().numpy().astype(np.complex128), Y) < 1e-5


@pytest.mark.parametrize("lx,ly,nu,nd", [(3, 2, 3, 3), (4, 2, 2, 3)])
def test_kblock_spectra_union_equals_full_sector(oracle, lx, ly, nu, nd):
    """At zero flux H commutes with translations: the block spectra together are the sector's."""
    say("kblock spectra", lx, ly, nu, nd)
    p = model(lx, ly)
    full = np.linalg.eigvalsh(oracle.dense_hamiltonian(params8(lx, ly, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0), nu, nd))
    ev, first = [], None
    for ky in range(ly):
        for kx in range(lx):
            B = api.KBlockOperator.from_params(p, nu, nd, api.MomentumSector(kx, ky, lx, ly), share=first)
            first = first or B
            d = B.dim()
            H = np.zeros((d, d), np.complex128)
            for j in range(d):
                e = np.zeros(d, np.complex128)
                e[j] = 1.0
                H[:, j] = B.apply(e)
            assert np.max(np.abs(H - H.conj().T)) < 1e-12
            ev.extend(np.linalg.eigvalsh(H))
    ev = np.sort(np.array(ev))
    assert len(ev) == len(full)
    assert np.max(np.abs(ev - full)) < 1e-9


def test_kblock_dims_sum_4x4():
    """Sum over the 16 momenta of the block dimensions is the sector dimension (every orbit
    survives at |G|/|Stab| momenta), at 4x4 (4,4) (dim 3,312,400)."""
    say("kblock dims 4x4 (4,4)")
    p = model(4, 4)
    tot, first = 0, None
    for ky in range(4):
        for kx in range(4):
            B = api.KBlockOperator.from_params(p, 4, 4, api.MomentumSector(kx, ky, 4, 4), share=first)
            first = first or B
            tot += B.dim()
            if (kx, ky) == (1, 2):
                # Hermiticity of a large block on random vectors: <x, H y> = <H x, y>
                rng = np.random.default_rng(4)
                x = rng.standard_normal(B.dim()) + 1j * rng.standard_normal(B.dim())
                y = rng.standard_normal(B.dim()) + 1j * rng.standard_normal(B.dim())
                a, b = np.vdot(x, B.apply(y)), np.vdot(B.apply(x), y)
                assert abs(a - b) < 1e-10 * abs(a)
    assert tot == 1820 * 1820


def test_kblock_split_paths_4x4_4_4(oracle):
    """4x4 (4,4) K=(1,2) block (about 2e5 rows): real bonds and D_up = D_dn = 1820 >= 128, so the
    apply splits (the down hops of the orbits whose up string no translation fixes go through the
    full sector's row pass: the gather row pass for R=1 complex128, the LDS-staged step row pass
    for the 128-byte elements of R=8 complex128 and R=16 complex64; the other terms through the
    block kernel). 1024 random rows of each column against the oracle's row-restricted orbit push
    (src/orbit_k_block_basis.cpp:151-220) at 1e-12 relative; the R=8 columns against the R=1
    applies at 1e-14 (the two row passes add the same down hops in other orders); complex64 at
    1e-5."""
    import torch

    say("kblock split 4x4 (4,4) K=(1,2)")
    B = kblock(4, 4, 4, 4, 1, 2)
    d = B.dim()
    reps = B.entries().rep
    rng = np.random.default_rng(17)
    R = 8
    X = rng.standard_normal((d, R)) + 1j * rng.standard_normal((d, R))
    xd = torch.tensor(X, dtype=torch.complex128, device="cuda").contiguous()
    yd = torch.empty_like(xd)
    B.apply_device(xd, yd, R=R, dtype=L.FTLM_C128)
    Y = yd.cpu().numpy()
    rows = np.sort(rng.choice(d, size=1024, replace=False))
    for r in range(R):
        y1 = B.apply(np.ascontiguousarray(X[:, r]))
        want = oracle.orbit_apply_rows(params8(4, 4, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0), 4, 4, 1, 2, reps,
                                       np.ascontiguousarray(X[:, r]), rows)
        e_or = np.max(np.abs(y1[rows] - want)) / np.max(np.abs(want))
        e_st = np.max(np.abs(Y[rows, r] - want)) / np.max(np.abs(want))
        e_12 = rel(Y[:, r], y1)
        say("  r", r, "R=1 vs oracle", e_or, "R=8 vs oracle", e_st, "R=8 vs R=1", e_12)
        assert e_or < 1e-12 and e_st < 1e-12 and e_12 < 1e-14
    X16 = rng.standard_normal((d, 16)) + 1j * rng.standard_normal((d, 16))
    x16 = torch.tensor(X16, dtype=torch.complex128, device="cuda").contiguous()
    y16 = torch.empty_like(x16)
    B.apply_device(x16, y16, R=16, dtype=L.FTLM_C128)
    x64 = x16.to(torch.complex64)
    y64 = torch.empty_like(x64)
    B.apply_device(x64, y64, R=16, dtype=L.FTLM_C64)
    e64 = rel(y64.cpu().numpy().astype(np.complex128), y16.cpu().numpy())
    say("  c64 R=16 vs c128", e64)
    assert e64 < 1e-5


def dense_of(B):
    d = B.dim()
    H = np.zeros((d, d), np.complex128)
    for j in range(d):
        e = np.zeros(d, np.complex128)
        e[j] = 1.0
        H[:, j] = B.apply(e)
    return H


@pytest.mark.parametrize("lx,ly,nu,nd,kx,ky", [(2, 2, 2, 2, 0, 0), (2, 2, 2, 2, 1, 1), (3, 2, 3, 3, 1, 0),
                                               (3, 2, 2, 1, 2, 1)])
def test_small_block_ed_vs_oracle(oracle, lx, ly, nu, nd, kx, ky):
    """exact_log_partition_small_complex: device ED of a k-block against the oracle's restatement
    on the same dense block (the capped Jacobi is deterministic given its input)."""
    say("ed kblock", lx, ly, nu, nd, kx, ky)
    B = kblock(lx, ly, nu, nd, kx, ky)
    H = dense_of(B)
    betas = np.array([0.1, 1.0, 5.0, 20.0])
    got, ev = api.exact_log_partition(B, betas, return_evals=True)
    want = np.array([oracle.exact_log_partition(H, b) for b in betas])
    w = np.linalg.eigvalsh(H)
    exact = np.array([np.max(-b * w) + np.log(np.sum(np.exp(-b * w - np.max(-b * w)))) for b in betas])
    # Beyond ~40 rows the reference's cap of max(5000, 150 m) rotations stops the Jacobi before it
    # converges, and the capped diagonal then depends on ulp-level differences of atan2/cos/sin
    # between the device and the host: 1e-9 relative; both stay equally close to the exact lnZ.
    assert np.max(np.abs(got - want) / np.maximum(1.0, np.abs(want))) < 1e-9, (got, want)
    assert np.max(np.abs(np.abs(got - exact) - np.abs(want - exact))) < 1e-9 * np.max(np.abs(exact))
    if 2 * B.dim() <= 40:  # converged well inside the rotation cap: the exact spectrum, twice
        assert np.max(np.abs(got - want)) < 1e-10
        assert np.max(np.abs(np.sort(ev) - np.sort(np.concatenate([w, w])))) < 1e-10
        for b, g in zip(betas, got):
            mx = np.max(-b * w)
            assert abs(g - (mx + np.log(np.sum(np.exp(-b * w - mx))))) < 1e-10


def test_small_sector_ed(oracle, golden):
    """Full sectors: dim 1 (-beta H00) and the 2x2 (2,2) sector of the 2x2 configuration."""
    say("ed sectors")
    m = golden["geom/c1_2x2/model"]
    p = api.HubbardParams(int(m[0]), int(m[1]), *map(float, m[2:]))
    op = api.HubbardOperator.from_params(p, 4, 4)
    assert op.dim() == 1
    h00 = op.apply(np.ones(1, np.complex128))[0].real
    assert api.exact_log_partition(op, 2.0) == -2.0 * h00
    op = api.HubbardOperator.from_params(p, 2, 2)
    H = oracle.dense_hamiltonian(params8(*m), 2, 2)
    got = api.exact_log_partition(op, np.array([0.5, 3.0]))
    want = [oracle.exact_log_partition(H, b) for b in (0.5, 3.0)]
    assert np.max(np.abs(got - want)) < 1e-9


@pytest.mark.parametrize("lx,ly,nu,nd,kx,ky,phi", [(4, 2, 3, 3, 1, 0, 0.05), (4, 2, 4, 4, 0, 0, 0.0)])
def test_kblock_gradient_vs_central_differences(lx, ly, nu, nd, kx, ky, phi):
    """d lnZ_K / d(t, t', U, V, phi_x) of one momentum block's FTLM estimator (same seeds,
    converged Gauss rule) against central differences of the block's lnZ with h = 1e-5."""
    say("kblock gradient FD", lx, ly, nu, nd, kx, ky, phi)
    base = dict(Lx=lx, Ly=ly, t=1.0, tp=-0.25, U=8.0, V=1.0, phi_x=phi, phi_y=0.0)
    K = api.MomentumSector(kx, ky, lx, ly)
    betas = np.array([0.5, 2.0, 8.0])
    par = api.FtlmParams(n_random=4, lanczos_steps=24, seed=9, quad=L.FTLM_QUAD_CONVERGED, clamp_M_to_dim=True)
    names = ("t", "tp", "U", "V", "phi_x")
    lnz, g = api.ftlm_log_partition_gradient(api.HubbardParams(**base), nu, nd, betas, par, params=names, K=K)
    B = api.KBlockOperator.from_params(api.HubbardParams(**base), nu, nd, K)
    assert np.max(np.abs(lnz - api.ftlm_log_partition_complex(B, betas, par))) < 1e-9
    h = 1e-5
    for name in names:
        vals = []
        for sgn in (+1, -1):
            q = dict(base)
            q[name] += sgn * h
            vals.append(api.ftlm_log_partition_complex(
                api.KBlockOperator.from_params(api.HubbardParams(**q), nu, nd, K), betas, par))
        fd = (vals[0] - vals[1]) / (2 * h)
        err = np.abs(fd - g[name]) / np.maximum(np.abs(fd), 1e-3)
        assert np.max(err) < 1e-6, (name, fd, g[name])


def test_kblock_vjp_matches_forward_mode():
    """Momentum block adjoint (term-split pass over the block's orbit terms, and per-operator
    parameter adjoints) against forward mode at phi = 0.  With flux the reference's block operator
    is not Hermitian (its Peierls phases sit on the wrapped bonds, which translations move): the
    adjoint sweep refuses it."""
    say("kblock vjp")
    p = model(4, 2, 0.0)
    K = api.MomentumSector(1, 0, 4, 2)
    betas = np.array([0.5, 2.0, 8.0])
    w = np.array([0.3, -1.1, 0.7])
    par = api.FtlmParams(n_random=4, lanczos_steps=24, seed=9, quad=L.FTLM_QUAD_CONVERGED, clamp_M_to_dim=True)
    names = ("t", "tp", "U", "V", "phi_x")
    lnz_f, g = api.ftlm_log_partition_gradient(p, 3, 3, betas, par, params=names, K=K)
    for mode in (L.FTLM_TS_FUSED, L.FTLM_TS_GENERIC):
        lnz, got = api.ftlm_log_partition_vjp(p, 3, 3, betas, w, par, params=names, K=K, checkpoint=5, mode=mode)
        assert np.max(np.abs(lnz - lnz_f)) < 1e-10
        for k in names:
            want = float(np.dot(w, g[k]))
            assert abs(got[k] - want) <= 1e-9 * max(abs(want), 1e-3), (mode, k, got[k], want)
    with pytest.raises(Exception, match="not Hermitian"):
        api.ftlm_log_partition_vjp(model(4, 2, 0.05), 3, 3, betas, w, par, params=names, K=K)


def test_symmetry_known_answers_on_device():
    """The reference's translation-symmetry known answers (test_symmetry_translation_stage1.cpp,
    test_symmetry_stage2.cpp, test_symmetry_stage3.cpp, test_momentum_partition_2x2.cpp) on the
    device primitives ftlm_hip_translate_states / _canonical_states / _bloch_phase."""
    import symmetry_kat as kat

    say("symmetry known answers on the device")

    def translate(lx, ly, u, d, dx, dy):
        out, sg = api.translate_states(lx, ly, np.array([(u << 16) | d], np.uint32), dx, dy)
        return int(out[0]), float(sg[0])

    def canonical(lx, ly, u, d):
        can, ss = api.canonical_states(lx, ly, np.array([(u << 16) | d], np.uint32))
        return int(can[0]), int(ss[0])

    def bloch(lx, ly, kx, ky, dx, dy):
        return api.translation_bloch_phase(api.MomentumSector(kx, ky, lx, ly), dx, dy)

    def sector(n, nu, nd):
        return api.FockBasis(n, nu, nd).masks()

    kat.check_translation_stage1(translate)
    kat.check_symmetry_stage2(translate, canonical)
    kat.check_symmetry_stage3(translate, canonical, bloch)
    kat.check_momentum_partition_2x2(translate, canonical, sector)
