"""GPU: reorthogonalised Lanczos (FTLM_REORTH_FULL / _PARTIAL) against what the mathematics fixes.

The reference never reorthogonalises, so nothing of it pins these modes; the checks are
- full Krylov space: with every Lanczos vector kept orthogonal, the recurrence breaks down once
  the space is exhausted and the Gauss rule integrates e^{-beta H} exactly, so per sample
  ln sum_k w_k e^{-beta lambda_k} = ln <r|e^{-beta H}|r>/<r|r> from a dense eigendecomposition
  of the oracle's H and the oracle's start vector (tolerance 1e-9 absolute), and every Ritz value
  is an eigenvalue (1e-9);
- no ghosts: below the space size, each isolated extreme eigenvalue appears once among the Ritz
  values (the plain recurrence duplicates the top one by step 240);
- the early steps, where the plain recurrence is still orthogonal to rounding, agree with it
  (alpha/beta within 1e-12 relative);
- partial (omega-monitored) against full: lnZ within 1e-9, complex64 storage against complex128
  within 1e-6; the graph-captured recurrence against the eager one.
"""

import sys

import numpy as np
import pytest

from pyoracle import params8

pytestmark = pytest.mark.gpu

api = pytest.importorskip("ftlm_extended_hubbard_amd.api")
from ftlm_extended_hubbard_amd import _lib as L  # noqa: E402

MODEL = dict(t=1.0, tp=-0.25, U=4.0, V=0.5, phi_x=0.1, phi_y=0.0)
BETAS = np.array([0.3, 1.0, 4.0])


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if api.device_count() < 1:
        pytest.skip("no HIP device")


def _op(Lx, Ly, nu, nd):
    p = api.HubbardParams(Lx, Ly, MODEL["t"], MODEL["tp"], MODEL["U"], MODEL["V"], MODEL["phi_x"], MODEL["phi_y"])
    return api.HubbardOperator.from_params(p, nu, nd)


def _p8(Lx, Ly):
    return params8(Lx, Ly, **MODEL)


def _ritz(op, R, M, seed, reorth, dtype=L.FTLM_C128, path=L.FTLM_PATH_PER_OP, tri=False):
    par = api.FtlmParams(n_random=R, lanczos_steps=M, seed=seed, start=L.FTLM_START_GAUSS_MT, dtype=dtype,
                         quad=L.FTLM_QUAD_CONVERGED, path=path, reorth=reorth)
    return api.ftlm_ritz([op], par, keep_tridiagonals=tri)


def _sample_logf(rd, r, betas):
    n = int(rd.used[0, r])
    lam, w = rd.ritz[0, r, :n], rd.weight[0, r, :n]
    x = -np.outer(betas, lam)
    m = x.max(axis=1, keepdims=True)
    return (m[:, 0] + np.log(np.sum(w * np.exp(x - m), axis=1)))


def _exact_logf(E, V, q, betas):
    c = np.abs(V.conj().T @ q) ** 2 / np.vdot(q, q).real
    x = -np.outer(betas, E)
    m = x.max(axis=1, keepdims=True)
    return m[:, 0] + np.log(np.sum(c * np.exp(x - m), axis=1))


def test_full_reorth_exhausts_krylov_space_exactly(oracle):
    Lx, Ly, nu, nd = 4, 2, 2, 1
    say("reorth full krylov 4x2", nu, nd)
    op = _op(Lx, Ly, nu, nd)
    dim = op.dim()
    H = oracle.dense_hamiltonian(_p8(Lx, Ly), nu, nd)
    assert np.allclose(H, H.conj().T, atol=1e-13)
    E, V = np.linalg.eigh(H)
    R, seed = 4, 5
    rd = _ritz(op, R, dim, seed, L.FTLM_REORTH_FULL)
    distinct = np.unique(np.round(E, 9))
    for r in range(R):
        n = int(rd.used[0, r])
        # breakdown no later than the number of distinct eigenvalues (plus rounding slack of one)
        assert n <= len(distinct) + 1, (n, len(distinct))
        q = oracle.gauss_start(dim, seed + r * 100003)
        got, want = _sample_logf(rd, r, BETAS), _exact_logf(E, V, q, BETAS)
        assert np.max(np.abs(got - want)) < 1e-9, (r, got, want)
        lam = rd.ritz[0, r, :n][rd.weight[0, r, :n] > 1e-10]
        assert np.max(np.min(np.abs(lam[:, None] - E[None, :]), axis=1)) < 1e-9


def _copies(rd, r, e, tol=1e-8):
    n = int(rd.used[0, r])
    return int(np.sum(np.abs(rd.ritz[0, r, :n] - e) < tol))


def test_reorth_removes_ghosts(oracle):
    Lx, Ly, nu, nd = 4, 2, 2, 2
    say("reorth ghosts 4x2", nu, nd)
    op = _op(Lx, Ly, nu, nd)
    H = oracle.dense_hamiltonian(_p8(Lx, Ly), nu, nd)
    E = np.linalg.eigvalsh(H)
    assert E[-1] - E[-2] > 1e-2 and E[1] - E[0] > 1e-3  # isolated extremes
    R, M, seed = 2, 240, 3
    plain = _ritz(op, R, M, seed, L.FTLM_REORTH_NONE)
    full = _ritz(op, R, M, seed, L.FTLM_REORTH_FULL)
    part = _ritz(op, R, M, seed, L.FTLM_REORTH_PARTIAL)
    for r in range(R):
        say("sample", r, "top copies plain/full/partial", _copies(plain, r, E[-1]), _copies(full, r, E[-1]),
            _copies(part, r, E[-1]), "ground copies", _copies(plain, r, E[0]), _copies(full, r, E[0]),
            _copies(part, r, E[0]))
        # the plain recurrence has lost orthogonality by M = 240: the top eigenvalue has ghosts
        assert _copies(plain, r, E[-1]) >= 2
        for rd in (full, part):
            assert _copies(rd, r, E[-1]) == 1 and _copies(rd, r, E[0]) == 1
            n = int(rd.used[0, r])
            lam = rd.ritz[0, r, :n]
            # below the space size the Ritz values interlace the spectrum: none leaves [E0, Emax]
            assert lam.min() > E[0] - 1e-9 and lam.max() < E[-1] + 1e-9


def test_reorth_early_steps_match_plain():
    say("reorth early steps 4x2 2 2")
    op = _op(4, 2, 2, 2)
    plain = _ritz(op, 3, 60, 11, L.FTLM_REORTH_NONE, tri=True)
    for mode in (L.FTLM_REORTH_FULL, L.FTLM_REORTH_PARTIAL):
        ro = _ritz(op, 3, 60, 11, mode, tri=True)
        k = 8
        for name in ("alpha", "beta"):
            a, b = getattr(ro, name)[0, :, :k], getattr(plain, name)[0, :, :k]
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), (mode, name)


def test_partial_matches_full_and_c64_matches_c128():
    say("reorth partial vs full, c64 vs c128, 4x2 2 2 and 4x3 3 3")
    for (Lx, Ly, nu, nd), M in (((4, 2, 2, 2), 200), ((4, 3, 3, 3), 120)):
        op = _op(Lx, Ly, nu, nd)
        dims = np.array([op.dim()], np.int64)
        R = 4
        out = {}
        for mode in (L.FTLM_REORTH_FULL, L.FTLM_REORTH_PARTIAL):
            for dt in (L.FTLM_C128, L.FTLM_C64):
                rd = _ritz(op, R, M, 21, mode, dtype=dt)
                out[mode, dt] = api.ftlm_thermo(rd, BETAS)["lnZ"][0]
        say(Lx, Ly, nu, nd, {k: v.tolist() for k, v in out.items()})
        full, part = out[L.FTLM_REORTH_FULL, L.FTLM_C128], out[L.FTLM_REORTH_PARTIAL, L.FTLM_C128]
        assert np.max(np.abs(full - part)) < 1e-9
        for mode in (L.FTLM_REORTH_FULL, L.FTLM_REORTH_PARTIAL):
            assert np.max(np.abs(out[mode, L.FTLM_C64] - out[mode, L.FTLM_C128])) < 1e-6
        assert dims[0] > 0


def test_reorth_graph_path_identical_to_eager():
    say("reorth graph vs eager 4x2 2 2")
    op = _op(4, 2, 2, 2)
    for mode in (L.FTLM_REORTH_FULL, L.FTLM_REORTH_PARTIAL):
        eager = _ritz(op, 3, 80, 7, mode, tri=True)
        graph = _ritz(op, 3, 80, 7, mode, path=L.FTLM_PATH_GRAPH, tri=True)
        assert np.array_equal(eager.alpha, graph.alpha) and np.array_equal(eager.beta, graph.beta)
        assert np.array_equal(eager.used, graph.used)


def test_reorth_refused_where_unsupported():
    say("reorth refusals")
    op = _op(4, 2, 2, 2)
    par = api.FtlmParams(n_random=2, lanczos_steps=10, path=L.FTLM_PATH_PERSISTENT, reorth=L.FTLM_REORTH_FULL)
    with pytest.raises(Exception, match="reorthogonalise"):
        api.ftlm_ritz([op], par)
    p = api.HubbardParams(4, 2, 1.0, -0.25, 4.0, 0.5, 0.0, 0.0)
    par = api.FtlmParams(n_random=2, lanczos_steps=10, start=L.FTLM_START_PHASE, reorth=L.FTLM_REORTH_FULL)
    with pytest.raises(Exception, match="unreorthogonalised"):
        api.ftlm_log_partition_gradient(p, 2, 2, BETAS, par)
