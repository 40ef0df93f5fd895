"""GPU test of the allocation-free thermodynamics entry point ftlm_hip_ftlm_thermo_noalloc (through the
C ABI): on identical inputs it returns exactly what ftlm_hip_ftlm_thermo returns (same kernels, so
bit-identical), on two calls in a row with the second smaller than the first, so that the
workspace left by the first call cannot leak into the second; a request beyond its 16 MB workspace
is refused with FTLM_EUNSUPPORTED."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

api = pytest.importorskip("ftlm_extended_hubbard_amd.api")
from ftlm_extended_hubbard_amd import _lib as L  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _device():
    if api.device_count() < 1:
        pytest.skip("no HIP device")


def ritz_data(rng, n_ops, R, M):
    ritz = np.sort(rng.uniform(-20.0, 40.0, (n_ops, R, M)), axis=2)
    w = rng.uniform(0.0, 1.0, (n_ops, R, M))
    used = rng.integers(0, M + 1, (n_ops, R)).astype(np.int32)
    used[:, 0] = M  # every operator keeps at least one sample
    for i in range(n_ops):
        for r in range(R):
            w[i, r, used[i, r]:] = 0.0
            s = w[i, r].sum()
            if s > 0:
                w[i, r] /= s
    dims = rng.integers(1, 5000, n_ops).astype(np.int64)
    N = rng.integers(0, 17, n_ops).astype(np.int32)
    return ritz, w, used, dims, N


def call(fn, ritz, w, used, dims, N, betas, mus):
    n_ops, R, M = ritz.shape
    nb, nmu = len(betas), len(mus)
    out = [np.full((n_ops, nb), np.nan), np.full((n_ops, nb), np.nan)] + \
          [np.full((nb, max(nmu, 1)), np.nan) for _ in range(3)]
    rc = fn(ritz.ctypes.data, w.ctypes.data, used.ctypes.data, n_ops, R, M, dims.ctypes.data,
            N.ctypes.data if nmu else None, betas.ctypes.data, nb, mus.ctypes.data if nmu else None, nmu,
            *[o.ctypes.data for o in out])
    return rc, out


@pytest.mark.parametrize("seed", [1, 2])
def test_noalloc_matches_thermo_twice_second_smaller(seed):
    lib = L.lib()
    rng = np.random.default_rng(seed)
    cases = [(81, 64, 80, 50, 41), (7, 3, 12, 4, 0)]  # the second call is smaller than the first
    for n_ops, R, M, nb, nmu in cases:
        print("[case] noalloc thermo", seed, n_ops, R, M, nb, nmu, flush=True)
        ritz, w, used, dims, N = ritz_data(rng, n_ops, R, M)
        betas = np.geomspace(0.2, 20.0, nb)
        mus = np.linspace(-2.0, 10.0, nmu)
        rc0, want = call(lib.ftlm_hip_ftlm_thermo, ritz, w, used, dims, N, betas, mus)
        rc1, got = call(lib.ftlm_hip_ftlm_thermo_noalloc, ritz, w, used, dims, N, betas, mus)
        assert rc0 == 0 and rc1 == 0, L.lib().ftlm_hip_last_error()
        for a, b in zip(got[:2] + (got[2:] if nmu else []), want[:2] + (want[2:] if nmu else [])):
            assert np.array_equal(a, b)
        assert np.all(np.isfinite(got[0]))


def test_noalloc_refuses_requests_beyond_its_workspace():
    lib = L.lib()
    rng = np.random.default_rng(3)
    ritz, w, used, dims, N = ritz_data(rng, 4, 512, 600)  # 2 x 1.2M doubles > 16 MB
    betas = np.array([1.0])
    rc, _ = call(lib.ftlm_hip_ftlm_thermo_noalloc, ritz, w, used, dims, N, betas, np.zeros(0))
    assert rc == L.FTLM_EUNSUPPORTED
    msg = lib.ftlm_hip_last_error().decode()
    assert "workspace" in msg, msg
    # the entry point still works afterwards
    ritz, w, used, dims, N = ritz_data(rng, 3, 2, 5)
    rc0, want = call(lib.ftlm_hip_ftlm_thermo, ritz, w, used, dims, N, betas, np.zeros(0))
    rc1, got = call(lib.ftlm_hip_ftlm_thermo_noalloc, ritz, w, used, dims, N, betas, np.zeros(0))
    assert rc0 == rc1 == 0
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
