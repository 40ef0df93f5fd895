"""GPU tests at the sizes of the workload configurations that the smaller parity tests do not reach:

- 4x4 (8,8) momentum blocks (all 16 K, phi = 0): the block dimensions sum to the sector dimension,
  2048 sampled rows of one block's H_K x match the oracle's row-restricted orbit push at 1e-12, and
  block Lanczos from v gives the same alpha/beta as full-sector Lanczos from the Bloch lift
  Phi_K v (H Phi_K = Phi_K H_K at zero flux), also on 4x4 (4,4);
- 4x3 (6,6) (dim 853,776) derivatives of lnZ: forward mode against central differences (h = 1e-5)
  and the adjoint sweep against forward mode; the adjoint on the 4x4 (8,8) K=(0,0) block;
- the device thermodynamic sums (per-operator lnZ and E, grand-canonical lnZ, <N>, E) on exact
  eigen-data of the 2x2 sectors (ED values and the published 2x2 n(mu) at beta = 20) and, on FTLM
  data of all 81 4x2 sectors, the identities E = -dlnZ/dbeta, E_gc = mu N_gc - dlnZ_gc/dbeta and
  N_gc = dlnZ_gc/dmu / beta;
- the 4x2 grand-canonical run at its own size (R = 64, M = min(dim, 80)): the persistent complex64
  launch against the per-sector complex128 path with the same starts.

Tolerances are stated in each test.
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


@pytest.fixture(scope="module", autouse=True)
def _device():
    if api.device_count() < 1:
        pytest.skip("no HIP device")


# ------------------------------------------------------------------ 4x4 (8,8) momentum blocks

C4 = dict(Lx=4, Ly=4, t=1.0, tp=-0.25, U=8.0, V=1.0, phi_x=0.0, phi_y=0.0)


def c4_params():
    return api.HubbardParams(**C4)


def test_c4_all_16_blocks_dims_and_sampled_rows(oracle):
    """All 16 momentum blocks of 4x4 (8,8): sum_K dim_K = 12870^2, representatives strictly
    increasing, and 2048 random rows of H_K x (complex128) on the largest block against the
    oracle's row-restricted orbit push (src/orbit_k_block_basis.cpp:151-220) at 1e-12 relative."""
    say("c4 dims + rows 4x4 (8,8)")
    p = c4_params()
    first = api.KBlockOperator.from_params(p, 8, 8, api.MomentumSector(0, 0, 4, 4))
    blocks = {(0, 0): first}
    for ky in range(4):
        for kx in range(4):
            if (kx, ky) != (0, 0):
                blocks[kx, ky] = api.KBlockOperator.from_params(p, 8, 8, api.MomentumSector(kx, ky, 4, 4),
                                                                share=first)
    dims = {k: b.dim() for k, b in blocks.items()}
    say("dims", dims)
    assert sum(dims.values()) == 12870 * 12870
    kx, ky = max(dims, key=dims.get)
    B = blocks[kx, ky]
    reps = B.entries().rep
    assert len(reps) == B.dim() and np.all(np.diff(reps.astype(np.int64)) > 0)
    rng = np.random.default_rng(41)
    x = rng.standard_normal(B.dim()) + 1j * rng.standard_normal(B.dim())
    y = B.apply(x)
    rows = np.sort(rng.choice(B.dim(), size=2048, replace=False))
    want = oracle.orbit_apply_rows(params8(**C4), 8, 8, kx, ky, reps, x, rows)
    err = np.max(np.abs(y[rows] - want)) / np.max(np.abs(want))
    say("K", (kx, ky), "dim", B.dim(), "row err", err)
    assert err < 1e-12


@pytest.mark.parametrize("nu,nd,kx,ky", [(4, 4, 1, 0), (8, 8, 1, 2)])
def test_c4_lifted_start_identity(oracle, nu, nd, kx, ky):
    """Block Lanczos from a normalised v_K and full-sector Lanczos from Phi_K v_K (the reference's
    Bloch sums, include/ftlm/symmetry/orbit_bloch_phi.hpp:35-51) produce the same alpha/beta: H
    commutes with translations at zero flux, so the full chain stays in the image of Phi_K.
    10 steps, complex128, 1e-10 relative to max |alpha|."""
    say("c4 lifted start", nu, nd, kx, ky)
    p = c4_params()
    B = api.KBlockOperator.from_params(p, nu, nd, api.MomentumSector(kx, ky, 4, 4))
    reps = B.entries().rep
    rng = np.random.default_rng(7)
    v = rng.standard_normal(B.dim()) + 1j * rng.standard_normal(B.dim())
    v /= np.linalg.norm(v)
    f = oracle.orbit_lift(4, 4, nu, nd, kx, ky, reps, v)
    assert abs(np.linalg.norm(f) - 1.0) < 1e-12
    M = 10
    ub, ab, bb = api.lanczos_tridiagonal(B, M, 1, q0=v)
    del B
    S = api.HubbardOperator.from_params(p, nu, nd)
    uf, af, bf = api.lanczos_tridiagonal(S, M, 1, q0=f)
    del f
    assert ub == uf == M
    scale = np.max(np.abs(ab))
    da, db = np.max(np.abs(ab - af)) / scale, np.max(np.abs(bb[:M - 1] - bf[:M - 1])) / scale
    say("alpha", ab[:3], "d_alpha", da, "d_beta", db)
    assert da < 1e-10 and db < 1e-10


# ------------------------------------------------------------------ 4x3 (6,6) derivatives

C5 = dict(Lx=4, Ly=3, t=1.0, tp=-0.25, U=8.0, V=1.0, phi_x=0.05, phi_y=0.0)
C5_MU = 8.0 / 2 + 2 * 1.0
C5_BETAS = np.array([0.5, 1.0, 2.0, 4.0, 8.0])
NAMES = ("t", "tp", "U", "V", "phi_x")


def test_c5_gradient_4x3_6_6_vs_central_differences():
    """d lnZ / d(t, t', U, V, phi_x) at 4x3 (6,6), dim 853,776, mu = U/2 + 2V, phi_x = 0.05,
    beta in {0.5, 1, 2, 4, 8} (R = 4 random-phase samples, M = 30, converged Gauss rule, same seeds
    on both sides): forward mode against central differences with h = 1e-5, within 1e-6 relative
    (plus 1e-9 absolute for the smallest phi derivatives)."""
    say("c5 gradient vs FD 4x3 (6,6)")
    par = api.FtlmParams(n_random=4, lanczos_steps=30, seed=5, start=L.FTLM_START_PHASE,
                         quad=L.FTLM_QUAD_CONVERGED)
    lnz, g = api.ftlm_log_partition_gradient(api.HubbardParams(**C5), 6, 6, C5_BETAS, par,
                                             params=NAMES + ("mu",), mu=C5_MU)
    op0 = api.HubbardOperator.from_params(api.HubbardParams(**C5), 6, 6)
    assert op0.dim() == 853776
    base = api.ftlm_log_partition_complex(op0, C5_BETAS, par)
    assert np.max(np.abs(lnz - (base + C5_BETAS * C5_MU * 12))) < 1e-9 * np.max(np.abs(lnz))
    h = 1e-5
    for name in NAMES:
        vals = []
        for sgn in (+1, -1):
            q = dict(C5)
            q[name] += sgn * h
            op = api.HubbardOperator.from_params(api.HubbardParams(**q), 6, 6)
            vals.append(api.ftlm_log_partition_complex(op, C5_BETAS, par))
        fd = (vals[0] - vals[1]) / (2 * h)
        bound = 1e-6 * np.abs(fd) + 1e-9
        say(name, "g", g[name], "fd", fd, "|g-fd|", np.abs(g[name] - fd), "bound", bound)
        assert np.all(np.abs(g[name] - fd) <= bound), name
    assert np.allclose(g["mu"], C5_BETAS * 12)


def test_c5_adjoint_vs_forward_4x3_6_6():
    """The adjoint sweep (sqrt(M) checkpoints, term-split parameter adjoints) against forward mode
    at 4x3 (6,6): sum_b w_b dlnZ(beta_b)/dtheta within 1e-9 relative."""
    say("c5 adjoint vs forward 4x3 (6,6)")
    par = api.FtlmParams(n_random=4, lanczos_steps=30, seed=5, start=L.FTLM_START_PHASE,
                         quad=L.FTLM_QUAD_CONVERGED)
    w = np.array([0.3, -1.1, 0.7, 0.2, -0.05])
    p = api.HubbardParams(**C5)
    lnz_f, g = api.ftlm_log_partition_gradient(p, 6, 6, C5_BETAS, par, params=NAMES)
    lnz, got = api.ftlm_log_partition_vjp(p, 6, 6, C5_BETAS, w, par, params=NAMES)
    assert np.max(np.abs(lnz - lnz_f)) < 1e-10 * np.max(np.abs(lnz_f))
    for k in NAMES:
        want = float(np.dot(w, g[k]))
        say(k, got[k], want)
        assert abs(got[k] - want) <= 1e-9 * max(abs(want), 1e-3), (k, got[k], want)


def test_c5_adjoint_vs_forward_4x4_kblock_K00():
    """The same check on the 4x4 (8,8) K = (0,0) momentum block (about 1.04e7 rows; phi = 0, so
    the block is Hermitian), R = 2, M = 12."""
    say("c5 adjoint vs forward 4x4 (8,8) K=(0,0)")
    p = c4_params()
    K = api.MomentumSector(0, 0, 4, 4)
    par = api.FtlmParams(n_random=2, lanczos_steps=12, seed=3, start=L.FTLM_START_PHASE,
                         quad=L.FTLM_QUAD_CONVERGED)
    w = np.array([1.0, 0.5, -0.25, 0.1, 0.05])
    names = ("t", "tp", "U", "V")
    lnz_f, g = api.ftlm_log_partition_gradient(p, 8, 8, C5_BETAS, par, params=names, K=K)
    lnz, got = api.ftlm_log_partition_vjp(p, 8, 8, C5_BETAS, w, par, params=names, K=K)
    assert np.max(np.abs(lnz - lnz_f)) < 1e-10 * np.max(np.abs(lnz_f))
    for k in names:
        want = float(np.dot(w, g[k]))
        say(k, got[k], want)
        assert abs(got[k] - want) <= 1e-9 * max(abs(want), 1e-3), (k, got[k], want)


# ------------------------------------------------------------------ device thermodynamic sums

def exact_ritz_data(oracle, p8, sectors):
    """Ritz data that make the FTLM sums exact: for each sector of dimension d, d 'samples' (the
    unit vectors) with every eigenvalue as a Ritz value and weight |psi_k(r)|^2."""
    evs = []
    for nu, nd in sectors:
        H = oracle.dense_hamiltonian(p8, nu, nd)
        evs.append(np.linalg.eigh(H))
    R = max(len(e[0]) for e in evs)
    n = len(sectors)
    rz, wt = np.zeros((n, R, R)), np.zeros((n, R, R))
    used = np.zeros((n, R), np.int32)
    dims = np.zeros(n, np.int64)
    for i, (lam, U) in enumerate(evs):
        d = len(lam)
        dims[i] = d
        for r in range(d):
            rz[i, r, :d] = lam
            wt[i, r, :d] = np.abs(U[r, :]) ** 2
            used[i, r] = d
    return api.RitzData(rz, wt, used, dims), [e[0] for e in evs]


def test_thermo_exact_data_2x2_vs_ed_and_published(oracle, golden):
    """ftlm_hip_ftlm_thermo on the exact eigen-data of all 25 2x2 sectors (t=1, t'=-0.35, U=5.75,
    V=0.9): per-sector lnZ and E equal ED's ln sum e^{-beta E} and <H> (1e-12 relative); the
    grand-canonical lnZ, <N> and E equal the ED mixture (1e-12); at beta = 20 <N>/4 reproduces the
    reference's published n(mu) (benchmarks/data/n_vs_mu_beta20.tsv, its printed precision)."""
    say("thermo exact 2x2")
    m = golden["geom/pub_2x2/model"]
    p8 = params8(*m)
    sectors = [(nu, nd) for nu in range(5) for nd in range(5)]
    rd, evs = exact_ritz_data(oracle, p8, sectors)
    N = np.array([nu + nd for nu, nd in sectors], np.int32)
    tab = golden["published/n_vs_mu_2x2_beta20"]
    betas = np.array([0.3, 2.0, 20.0])
    mus = tab[:, 0]
    th = api.ftlm_thermo(rd, betas, N=N, mus=mus)
    lz_ed = np.array([[np.logaddexp.reduce(-b * e) for b in betas] for e in evs])
    e_ed = np.array([[np.sum(e * np.exp(-b * e - np.max(-b * e))) / np.sum(np.exp(-b * e - np.max(-b * e)))
                      for b in betas] for e in evs])
    assert np.max(np.abs(th["lnZ"] - lz_ed) / np.maximum(1.0, np.abs(lz_ed))) < 1e-12
    assert np.max(np.abs(th["E"] - e_ed) / np.maximum(1.0, np.abs(e_ed))) < 1e-12
    for ib, b in enumerate(betas):
        x = lz_ed[:, ib][:, None] + b * np.outer(N, mus)  # [sector, mu]
        lg = np.logaddexp.reduce(x, axis=0)
        P = np.exp(x - lg)
        assert np.max(np.abs(th["lnZ_gc"][ib] - lg) / np.maximum(1.0, np.abs(lg))) < 1e-12
        assert np.max(np.abs(th["N_gc"][ib] - N @ P)) < 1e-12 * 8
        assert np.max(np.abs(th["E_gc"][ib] - e_ed[:, ib] @ P) / np.maximum(1.0, np.abs(e_ed[:, ib] @ P))) < 1e-11
    n20 = th["N_gc"][2] / 4
    assert np.allclose(n20, tab[:, 1], rtol=2e-5, atol=1e-12)
    # the oracle's restatement of the reference's mixing on the same per-sector lnZ agrees
    assert np.max(np.abs(oracle.gc_density(th["lnZ"][:, 2].copy(), N, 20.0, mus, 4) - n20)) < 1e-12


def test_thermo_identities_on_ftlm_data_4x2():
    """On FTLM Ritz data of all 81 4x2 sectors (R=4, M=min(dim,30), converged Gauss rule), the
    device sums obey, at every beta and mu (central differences, step 1e-6 relative):
    E_op = -dlnZ_op/dbeta, E_gc = mu N_gc - dlnZ_gc/dbeta, N_gc = dlnZ_gc/dmu / beta (1e-6)."""
    say("thermo identities 4x2")
    p = api.HubbardParams(4, 2, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0)
    ops, N = [], []
    for nu in range(9):
        for nd in range(9):
            ops.append(api.HubbardOperator.from_params(p, nu, nd))
            N.append(nu + nd)
    par = api.FtlmParams(n_random=4, lanczos_steps=30, seed=2, start=L.FTLM_START_PHASE,
                         quad=L.FTLM_QUAD_CONVERGED, clamp_M_to_dim=True)
    rd = api.ftlm_ritz(ops, par)
    betas = np.array([0.2, 1.0, 5.0, 20.0])
    mus = np.array([-1.0, 2.0, 4.5, 9.0])
    th = api.ftlm_thermo(rd, betas, N=N, mus=mus)
    eps = 1e-6
    hp = api.ftlm_thermo(rd, betas * (1 + eps), N=N, mus=mus / (1 + eps))  # beta*mu fixed
    hm = api.ftlm_thermo(rd, betas * (1 - eps), N=N, mus=mus / (1 - eps))
    dlz_op = (hp["lnZ"] - hm["lnZ"]) / (2 * eps * betas)
    assert np.max(np.abs(th["E"] + dlz_op) / np.maximum(1.0, np.abs(th["E"]))) < 1e-6
    # at fixed beta*mu, d lnZ_gc / dbeta = -E_gc
    dlz_gc = (hp["lnZ_gc"] - hm["lnZ_gc"]) / (2 * eps * betas[:, None])
    assert np.max(np.abs(th["E_gc"] + dlz_gc) / np.maximum(1.0, np.abs(th["E_gc"]))) < 1e-6
    # at fixed beta: dlnZ_gc/dmu = beta N_gc and dlnZ_gc/dbeta = mu N_gc - E_gc
    for ib, b in enumerate(betas):
        one = np.array([b])
        up = api.ftlm_thermo(rd, one, N=N, mus=mus + 1e-6)["lnZ_gc"][0]
        dn = api.ftlm_thermo(rd, one, N=N, mus=mus - 1e-6)["lnZ_gc"][0]
        assert np.max(np.abs((up - dn) / 2e-6 / b - th["N_gc"][ib])) < 1e-6 * 16
        bp = api.ftlm_thermo(rd, one * (1 + eps), N=N, mus=mus)["lnZ_gc"][0]
        bm = api.ftlm_thermo(rd, one * (1 - eps), N=N, mus=mus)["lnZ_gc"][0]
        want = mus * th["N_gc"][ib] - th["E_gc"][ib]
        assert np.max(np.abs((bp - bm) / (2 * eps * b) - want) / np.maximum(1.0, np.abs(want))) < 1e-6


def test_c2_complex64_persistent_vs_complex128_at_config():
    """The 4x2 grand-canonical configuration at its own size: 81 sectors, R = 64 random-phase
    starts, M = min(dim, 80), 50 temperatures in [0.05, 5].  The complex64 persistent launch (fp64
    dots, alpha/beta) against the per-sector complex128 path with the same starts.  complex64
    storage rounds every Krylov vector to 24 bits, which moves the Ritz values by ~1e-7 |E|; lnZ
    then moves by beta times that.  Bound: |dlnZ(beta)| <= max(1e-6, 2e-7 beta) per sector, and
    |d<N>| <= 1e-6 on the 41 x 50 (mu, T) grid."""
    say("c2 c64 vs c128 at R=64, M=80")
    p = api.HubbardParams(4, 2, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0)
    ops, N = [], []
    for nu in range(9):
        for nd in range(9):
            ops.append(api.HubbardOperator.from_params(p, nu, nd))
            N.append(nu + nd)
    dims = np.array([o.dim() for o in ops])
    betas = 1.0 / np.logspace(np.log10(0.05), np.log10(5.0), 50)
    mus = np.linspace(-2.0, 10.0, 41)
    out = {}
    for name, dt, path in (("c64", L.FTLM_C64, L.FTLM_PATH_PERSISTENT), ("c128", L.FTLM_C128, L.FTLM_PATH_PER_OP)):
        par = api.FtlmParams(n_random=64, lanczos_steps=80, seed=1, start=L.FTLM_START_PHASE, dtype=dt,
                             quad=L.FTLM_QUAD_CONVERGED, clamp_M_to_dim=True, path=path)
        out[name] = api.ftlm_thermo(api.ftlm_ritz(ops, par), betas, N=N, mus=mus)
    dl = np.abs(out["c64"]["lnZ"] - out["c128"]["lnZ"])[dims > 1]  # [sector, beta]
    worst = dl.max(axis=0)
    bound = np.maximum(1e-6, 2e-7 * betas)
    say("max |dlnZ| at beta", [(round(b, 2), float(w)) for b, w in zip(betas[::7], worst[::7])])
    assert np.all(worst <= bound), (worst / bound).max()
    dN = np.abs(out["c64"]["N_gc"] - out["c128"]["N_gc"]).max()
    say("max |d<N>|", dN)
    assert dN <= 1e-6
