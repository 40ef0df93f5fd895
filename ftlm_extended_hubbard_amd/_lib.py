"""ctypes binding of libftlm_hip.so (the C ABI declared in include/ftlm_hip.h).

Loading fails loudly: there is no CPU fallback anywhere in this package.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libftlm_hip.so")

FTLM_OK, FTLM_EINVAL, FTLM_ERANGE, FTLM_EHIP, FTLM_ENOMEM, FTLM_EUNSUPPORTED = range(6)
FTLM_C128, FTLM_C64 = 0, 1
FTLM_START_GAUSS_MT, FTLM_START_PHASE = 0, 1
FTLM_QUAD_REFERENCE, FTLM_QUAD_CONVERGED = 0, 1
FTLM_PATH_AUTO, FTLM_PATH_PER_OP, FTLM_PATH_PERSISTENT, FTLM_PATH_GRAPH, FTLM_PATH_STACK = 0, 1, 2, 3, 4
FTLM_REORTH_NONE, FTLM_REORTH_FULL, FTLM_REORTH_PARTIAL = 0, 1, 2
FTLM_TS_AUTO, FTLM_TS_GENERIC, FTLM_TS_FUSED = 0, 1, 2


class FtlmHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ftlm_hip code {code}] {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("Lx", C.c_int), ("Ly", C.c_int), ("t", C.c_double), ("tp", C.c_double), ("U", C.c_double),
                ("V", C.c_double), ("phi_x", C.c_double), ("phi_y", C.c_double)]


class Hop(C.Structure):
    _fields_ = [("from_", C.c_int), ("to", C.c_int), ("spin", C.c_int), ("re", C.c_double), ("im", C.c_double)]


class Pair(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int)]


class HvShape(C.Structure):
    _fields_ = [("rb", C.c_int), ("bpx", C.c_int), ("wlog", C.c_int), ("grid", C.c_int), ("ms", C.c_double),
                ("ms_default", C.c_double), ("n_tried", C.c_int)]


class FtlmOpts(C.Structure):
    _fields_ = [("R", C.c_int), ("M", C.c_int), ("seed", C.c_uint), ("start", C.c_int), ("dtype", C.c_int),
                ("quad", C.c_int), ("clamp_M_to_dim", C.c_int), ("real_variant", C.c_int), ("path", C.c_int),
                ("reorth", C.c_int)]


# C struct layout (ftlm_hop_t is 32 bytes: three ints, 4 bytes of padding, two doubles)
# void (*)(const double* x, double* y, void* ctx) -- ftlm_host_apply_fn
HOST_APPLY_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)

HOP_DTYPE = np.dtype([("from", np.int32), ("to", np.int32), ("spin", np.int32), ("re", np.float64),
                      ("im", np.float64)], align=True)
PAIR_DTYPE = np.dtype([("i", np.int32), ("j", np.int32)], align=True)
assert HOP_DTYPE.itemsize == C.sizeof(Hop) and PAIR_DTYPE.itemsize == C.sizeof(Pair)

_lib = None


def lib() -> C.CDLL:
    """Load libftlm_hip.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ftlm_extended_hubbard_amd.build`")
    # PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1.  Two HIP runtimes in
    # one process cannot share the device, so load torch's first when it is installed: the
    # dynamic linker then binds this library to that same runtime (identical SONAMEs).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i64p, dp, ip = C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.ftlm_hip_version.restype = C.c_char_p
    L.ftlm_hip_last_error.restype = C.c_char_p
    L.ftlm_hip_device_count.argtypes = [ip]
    L.ftlm_hip_nsites.argtypes = [C.POINTER(Params)]
    L.ftlm_hip_build_geometry.argtypes = [C.POINTER(Params), vp, C.c_int, ip, vp, C.c_int, ip]
    L.ftlm_hip_sector_create.argtypes = [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, C.c_double,
                                         C.c_double, C.POINTER(vp)]
    L.ftlm_hip_op_destroy.argtypes = [vp]
    L.ftlm_hip_op_dim.argtypes = [vp, i64p, i64p, i64p]
    L.ftlm_hip_sector_masks.argtypes = [vp, C.c_int64, C.c_int64, vp, vp]
    L.ftlm_hip_index_of.argtypes = [vp, vp, vp, C.c_int64, vp]
    L.ftlm_hip_apply.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp]
    L.ftlm_hip_apply_pass_times.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp, vp]
    L.ftlm_hip_autotune.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, C.POINTER(HvShape)]
    L.ftlm_hip_set_shape.argtypes = [vp, C.c_int, C.c_int, C.POINTER(HvShape)]
    L.ftlm_hip_get_shape.argtypes = [vp, C.c_int, C.c_int, C.POINTER(HvShape)]
    L.ftlm_hip_apply_host.argtypes = [vp, vp, vp]
    L.ftlm_hip_lanczos.argtypes = [vp, C.c_int, C.c_uint, vp, vp, vp, ip, ip]
    L.ftlm_hip_lanczos_extrema.argtypes = [vp, C.c_int, C.c_uint, dp, dp, ip]
    L.ftlm_hip_tridiagonal_extrema.argtypes = [vp, C.c_int, vp, C.c_int, dp, dp]
    L.ftlm_hip_ftlm_ritz.argtypes = [vp, C.c_int, C.POINTER(FtlmOpts), vp, vp, vp, vp, vp, vp]
    L.ftlm_hip_ftlm_thermo.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp, C.c_int,
                                       vp, vp, vp, vp, vp]
    L.ftlm_hip_ftlm_thermo_noalloc.argtypes = L.ftlm_hip_ftlm_thermo.argtypes
    L.ftlm_hip_ftlm_log_partition.argtypes = [vp, vp, C.c_int, C.POINTER(FtlmOpts), vp]
    L.ftlm_hip_quadrature_ritz.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, vp]
    L.ftlm_hip_krylov_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, C.POINTER(vp)]
    L.ftlm_hip_krylov_run.argtypes = [vp, C.c_int, vp]
    L.ftlm_hip_krylov_run_graph.argtypes = L.ftlm_hip_krylov_run.argtypes
    L.ftlm_hip_krylov_result.argtypes = [vp, vp, vp, vp, vp]
    L.ftlm_hip_krylov_bytes.argtypes = [vp, i64p]
    L.ftlm_hip_krylov_destroy.argtypes = [vp]
    L.ftlm_hip_build_geometry_derivative.argtypes = [C.POINTER(Params), C.c_int, vp, C.c_int, ip]
    L.ftlm_hip_ftlm_gradient.argtypes = [vp, vp, C.c_int, C.POINTER(FtlmOpts), vp, C.c_int, vp, vp]
    L.ftlm_hip_ftlm_vjp.argtypes = [vp, vp, C.c_int, C.POINTER(FtlmOpts), vp, C.c_int, vp, C.c_int, C.c_int, vp, vp]
    L.ftlm_hip_dense_jacobi.argtypes = [vp, C.c_int, C.c_int, C.c_double, vp, vp, vp]
    L.ftlm_hip_ftlm_observables.argtypes = [vp, vp, C.c_int, C.POINTER(FtlmOpts), vp, C.c_int, vp, vp]
    L.ftlm_hip_ftlm_tape_forward.argtypes = [vp, vp, C.c_int, C.POINTER(FtlmOpts), vp, C.c_int, C.c_int, C.c_int, vp,
                                             C.POINTER(vp)]
    L.ftlm_hip_ftlm_tape_backward.argtypes = [vp, vp, vp]
    L.ftlm_hip_ftlm_tape_destroy.argtypes = [vp]
    L.ftlm_hip_exact_log_partition.argtypes = [vp, vp, C.c_int, vp, vp]
    L.ftlm_hip_callback_op_create.argtypes = [C.c_int64, HOST_APPLY_FN, vp, C.POINTER(vp)]
    L.ftlm_hip_translate_states.argtypes = [C.c_int, C.c_int, vp, C.c_int64, C.c_int, C.c_int, vp, vp]
    L.ftlm_hip_canonical_states.argtypes = [C.c_int, C.c_int, vp, C.c_int64, vp, vp]
    L.ftlm_hip_bloch_phase.argtypes = [C.c_int] * 6 + [dp, dp]
    L.ftlm_hip_kblock_create.argtypes = [C.c_int] * 6 + [C.c_double, vp, C.c_int, vp, C.c_int, C.c_double,
                                                         C.c_double, vp, C.POINTER(vp)]
    L.ftlm_hip_kblock_info.argtypes = [vp, ip, ip, i64p, i64p]
    L.ftlm_hip_kgram_create.argtypes = [vp, vp, C.POINTER(C.c_void_p)]
    L.ftlm_hip_kblock_entries.argtypes = [vp, C.c_int64, C.c_int64, vp, vp, vp, vp]
    L.ftlm_hip_kblock_index_of_rep.argtypes = [vp, vp, C.c_int64, vp]
    L.ftlm_hip_kblock_map.argtypes = [vp, vp, C.c_int64, vp, vp, vp, vp]
    L.ftlm_hip_kstack_create.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.ftlm_hip_kstack_scatter.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, vp]
    L.ftlm_hip_kstack_gather.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, vp]
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != FTLM_OK:
        msg = lib().ftlm_hip_last_error().decode(errors="replace")
        if rc in (FTLM_EINVAL, FTLM_EUNSUPPORTED):
            raise ValueError(f"[ftlm_hip code {rc}] {msg}")
        if rc == FTLM_ERANGE:
            raise IndexError(f"[ftlm_hip code {rc}] {msg}")
        if rc == FTLM_ENOMEM:
            raise MemoryError(f"[ftlm_hip code {rc}] {msg}")
        raise FtlmHipError(rc, msg)


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def exported_symbols() -> list[str]:
    """Names of every entry point declared in include/ftlm_hip.h."""
    import re

    hdr = os.path.join(os.path.dirname(PKG), "include", "ftlm_hip.h")
    with open(hdr) as f:
        return sorted(set(re.findall(r"\b(ftlm_hip_[a-z_0-9]+)\s*\(", f.read())))
