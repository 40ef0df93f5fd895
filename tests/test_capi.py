"""CPU-side checks of the C ABI library: it loads, exports every declared entry point, and its host
logic (geometry, argument validation, error mapping) behaves as the reference.  No kernel launches."""

import ctypes as C
import math
import subprocess

import numpy as np
import pytest

from ftlm_extended_hubbard_amd import _lib as L
from ftlm_extended_hubbard_amd import api


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = L.exported_symbols()
    assert len(declared) >= 19
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(lib, s)


def test_version_and_nsites():
    assert api.version_string().startswith("0.1.0")
    assert api.nsites(api.HubbardParams(Lx=4, Ly=4)) == 16  # tests/smoke.cpp:5-10


@pytest.mark.parametrize("name", ["c1_2x2", "pub_2x2", "flux_2x2", "c2_4x2", "flux_3x2", "c5_4x3", "c3_4x4"])
def test_geometry_matches_reference(golden, name):
    m = golden[f"geom/{name}/model"]
    p = api.HubbardParams(int(m[0]), int(m[1]), *map(float, m[2:]))
    hops, pairs = api.build_hubbard_geometry(p)
    g = lambda k: golden[f"geom/{name}/{k}"]
    assert np.array_equal(hops["from"], g("hops_from"))
    assert np.array_equal(hops["to"], g("hops_to"))
    assert np.array_equal(hops["spin"], g("hops_spin"))
    assert np.array_equal(hops["re"] + 1j * hops["im"], g("hops_coeff"))
    assert np.array_equal(np.stack([pairs["i"], pairs["j"]], 1), g("pairs"))


def test_geometry_capacity_error():
    p = L.Params(4, 4, 1.0, -0.25, 0, 0, 0, 0)
    nh, np_ = C.c_int(0), C.c_int(0)
    hops = np.zeros(10, L.HOP_DTYPE)
    pairs = np.zeros(64, L.PAIR_DTYPE)
    rc = L.lib().ftlm_hip_build_geometry(C.byref(p), hops.ctypes.data, 10, C.byref(nh), pairs.ctypes.data, 64,
                                         C.byref(np_))
    assert rc == L.FTLM_ERANGE and nh.value == 384 and np_.value == 32


@pytest.mark.parametrize("n,nu,nd", [(0, 0, 0), (33, 1, 1), (4, 5, 0), (4, 0, -1)])
def test_invalid_sector_is_invalid_argument(n, nu, nd):
    # src/fock_basis.cpp:19-24 throws std::invalid_argument before touching any device memory
    h = C.c_void_p()
    rc = L.lib().ftlm_hip_sector_create(n, nu, nd, None, 0, None, 0, 0.0, 0.0, C.byref(h))
    assert rc == L.FTLM_EINVAL and not h.value
    assert b"FockBasis" in L.lib().ftlm_hip_last_error()
    with pytest.raises(ValueError):
        L.check(rc)


def test_bad_options_rejected_without_device():
    opts = L.FtlmOpts(4, 10, 1, 7, 0, 0, 0, 0)  # start = 7 is invalid
    rz = np.zeros(40)
    us = np.zeros(4, np.int32)
    arr = (C.c_void_p * 1)(None)
    rc = L.lib().ftlm_hip_ftlm_ritz(arr, 1, C.byref(opts), rz.ctypes.data, rz.ctypes.data, us.ctypes.data, None,
                                    None, None)
    assert rc == L.FTLM_EINVAL


def test_bad_reorth_rejected_without_device():
    opts = L.FtlmOpts(4, 10, 1, 0, 0, 0, 0, 0, 0, 3)  # reorth = 3 is no FTLM_REORTH_* mode
    rz = np.zeros(40)
    us = np.zeros(4, np.int32)
    arr = (C.c_void_p * 1)(None)
    rc = L.lib().ftlm_hip_ftlm_ritz(arr, 1, C.byref(opts), rz.ctypes.data, rz.ctypes.data, us.ctypes.data, None,
                                    None, None)
    assert rc == L.FTLM_EINVAL
    assert b"reorth" in L.lib().ftlm_hip_last_error()


def test_tridiagonal_extrema_empty_is_zero():
    lo, hi = api.tridiagonal_extrema(np.zeros(0), np.zeros(0), 0)
    assert (lo, hi) == (0.0, 0.0)
    assert math.isfinite(lo)


def _merged(hops, scale=1.0):
    import collections

    m = collections.defaultdict(complex)
    for e in hops:
        m[(int(e["from"]), int(e["to"]), int(e["spin"]))] += scale * (e["re"] + 1j * e["im"])
    return m


@pytest.mark.parametrize("model", [(4, 2, 1.0, -0.25, 8.0, 1.0, 0.1, 0.0), (4, 4, 0.7, -0.3, 8.0, 1.5, 0.2, -0.15),
                                   (2, 2, 1.0, -0.2, 4.0, 0.5, 0.31, -0.27), (4, 3, 1.0, -0.25, 8.0, 1.0, 0.05, 0.0)])
def test_geometry_derivative(model):
    """dH/dtheta hop lists: c = t dc/dt + t' dc/dt' on every merged bond (the amplitudes are linear
    in (t, t')), and dc/dphi equals a central difference of the geometry."""
    from ftlm_extended_hubbard_amd import api

    p = api.HubbardParams(*model)
    hops, _ = api.build_hubbard_geometry(p)
    a, b, c = _merged(hops), _merged(api.geometry_derivative(p, "t"), p.t), _merged(api.geometry_derivative(p, "tp"), p.tp)
    for k in set(a) | set(b) | set(c):
        assert abs(a[k] - b[k] - c[k]) < 1e-14, k
    h = 1e-6
    for which, idx in (("phi_x", 6), ("phi_y", 7)):
        mp, mm = list(model), list(model)
        mp[idx] += h
        mm[idx] -= h
        fp = _merged(api.build_hubbard_geometry(api.HubbardParams(*mp))[0])
        fm = _merged(api.build_hubbard_geometry(api.HubbardParams(*mm))[0])
        d = _merged(api.geometry_derivative(p, which))
        for k in set(fp) | set(fm) | set(d):
            assert abs((fp[k] - fm[k]) / (2 * h) - d[k]) < 1e-8, (which, k)


def test_autograd_module_surface():
    """The torch.autograd wrapper imports without a device and checks theta's shape first."""
    import torch

    from ftlm_extended_hubbard_amd import autograd as AG

    assert AG.THETA == ("t", "tp", "U", "V", "phi_x", "phi_y", "mu")
    with pytest.raises(ValueError, match="7 entries"):
        AG.ftlm_log_partition(torch.zeros(3, dtype=torch.float64), AG.FtlmModel(2, 2, 1, 1), [1.0],
                              AG.api.FtlmParams())
