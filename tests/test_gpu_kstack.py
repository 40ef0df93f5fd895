"""GPU parity of the momentum stack (all momentum blocks of one orbit graph in one apply and one
recurrence group) against the per-block operators it replaces, through the C ABI.

- apply: every block's slots of the stacked y equal that block's own apply (complex128: within
  1e-14 relative; on 4x4 sectors with real bonds both applies add the down hops of the orbits
  whose up string has no fixing translation first, through the full sector's row pass -- the
  LDS-staged step pass for the stack's wide elements, the gather pass for a block's narrow ones --
  and the other terms after them, so the two agree to rounding; complex64: 1e-6), and the slots
  of blocks an orbit is not a row of are 0, on 3x2, 4x2 (with a flux) and 4x4 (4,4) (16 blocks),
  and on two blocks of the 4x4 (8,8) stack of all 16;
- FTLM: ftlm_ritz through the stack (FTLM_PATH_STACK, and AUTO, which picks it) against the
  per-block path with the same random-phase starts: used steps equal, alpha/beta within 1e-10
  relative, Ritz values within 1e-9 and lnZ within 1e-10 (complex128; complex64 lnZ 1e-5), with
  sample counts that leave a partial recurrence group and M at most a quarter of the smallest block
  (closer to the block size, lost orthogonality amplifies the last-bit differences of the dot sums); on 4x4 (8,8), all 16 blocks at R = 2,
  M = 12;
- a Krylov session on the stack against per-block sessions (alpha/beta within 1e-10).
"""

import sys

import numpy as np
import pytest

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


def all_blocks(lx, ly, nu, nd, phi=0.0, U=8.0, V=1.0):
    p = api.HubbardParams(lx, ly, 1.0, -0.25, U, V, phi, 0.0)
    first = api.KBlockOperator.from_params(p, nu, nd, api.MomentumSector(0, 0, lx, ly))
    out = [first]
    for ky in range(ly):
        for kx in range(lx):
            if (kx, ky) != (0, 0):
                out.append(api.KBlockOperator.from_params(p, nu, nd, api.MomentumSector(kx, ky, lx, ly), share=first))
    return [b for b in out if b.dim() > 0]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def check_stack_apply(blocks, rk, dtype, seed, only=None):
    import torch

    ct = torch.complex128 if dtype == L.FTLM_C128 else torch.complex64
    S = api.KStackOperator.from_blocks(blocks)
    nk, n = len(blocks), S.dim()
    W = nk * rk
    g = torch.Generator(device="cuda").manual_seed(seed)
    xb = [torch.randn(b.dim() * rk, dtype=ct, device="cuda", generator=g) for b in blocks]
    xs = torch.empty(n * W, dtype=ct, device="cuda")
    for j in range(nk):
        S.scatter(j, xb[j], xs, rk=rk, dtype=dtype)
    ys = torch.full_like(xs, complex(7.0, 7.0))
    S.apply_device(xs, ys, R=W, dtype=dtype)
    # the stacked y rebuilt from the per-block applies: equal on the member slots, 0 elsewhere
    want = torch.empty_like(ys)
    worst = 0.0
    for j in (range(nk) if only is None else only):
        yb = torch.empty_like(xb[j])
        blocks[j].apply_device(xb[j], yb, R=rk, dtype=dtype)
        got = torch.empty_like(yb)
        S.gather(j, ys, got, rk=rk, dtype=dtype)
        worst = max(worst, rel(got.cpu().numpy().astype(np.complex128), yb.cpu().numpy().astype(np.complex128)))
        S.scatter(j, yb, want, rk=rk, dtype=dtype)
    if only is None:  # slots of blocks an orbit is not a row of: written as 0
        assert bool((ys[want == 0] == 0).all())
    return worst


@pytest.mark.parametrize("lx,ly,nu,nd,phi", [(3, 2, 3, 3, 0.0), (4, 2, 4, 4, 0.0), (4, 2, 2, 3, 0.05),
                                             (4, 4, 4, 4, 0.0)])
def test_stack_apply_equals_per_block(lx, ly, nu, nd, phi):
    blocks = all_blocks(lx, ly, nu, nd, phi)
    nk = len(blocks)
    say("stack apply", lx, ly, nu, nd, phi, "blocks", nk, "dims", [b.dim() for b in blocks])
    for rk in sorted({1, 2, 32 // nk}):
        if nk * rk > 32:
            continue
        e128 = check_stack_apply(blocks, rk, L.FTLM_C128, 11 + rk)
        e64 = check_stack_apply(blocks, rk, L.FTLM_C64, 23 + rk)
        say("  rk", rk, "c128 rel", e128, "c64 rel", e64)
        assert e128 <= 1e-14
        assert e64 <= 1e-6


def test_stack_apply_4x4_8_8_two_blocks():
    """The 16-block stack of 4x4 (8,8) (10.4 M orbits), rk = 1 complex128 and rk = 2 complex64:
    blocks K=(0,0) and K=(1,2) of the stacked y against their own applies."""
    blocks = all_blocks(4, 4, 8, 8)
    assert len(blocks) == 16 and sum(b.dim() for b in blocks) == 12870 * 12870
    order = [(kx, ky) for ky in range(4) for kx in range(4)]
    order.remove((0, 0))
    pick = [0, 1 + order.index((1, 2))]
    e128 = check_stack_apply(blocks, 1, L.FTLM_C128, 5, only=pick)
    e64 = check_stack_apply(blocks, 2, L.FTLM_C64, 6, only=pick)
    say("4x4 (8,8) stack: c128 rel", e128, "c64 rel", e64)
    assert e128 <= 1e-14 and e64 <= 1e-6


def ritz_pair(blocks, R, M, dtype, path, seed=3):
    par = api.FtlmParams(n_random=R, lanczos_steps=M, seed=seed, start=L.FTLM_START_PHASE, dtype=dtype,
                         quad=L.FTLM_QUAD_CONVERGED, path=L.FTLM_PATH_PER_OP)
    ref = api.ftlm_ritz(blocks, par, keep_tridiagonals=True)
    par.path = path
    got = api.ftlm_ritz(blocks, par, keep_tridiagonals=True)
    return ref, got


def lnz_of(rd, betas):
    return api.ftlm_thermo(rd, betas)["lnZ"]


@pytest.mark.parametrize("lx,ly,nu,nd,R,path", [(4, 2, 4, 4, 5, L.FTLM_PATH_STACK), (3, 2, 3, 3, 7, L.FTLM_PATH_AUTO),
                                                (4, 4, 4, 4, 3, L.FTLM_PATH_STACK)])
def test_stack_ftlm_matches_per_block(lx, ly, nu, nd, R, path):
    blocks = all_blocks(lx, ly, nu, nd)
    M = min(24, min(b.dim() for b in blocks) // 4)  # well inside the block: no ghost amplification
    betas = np.array([0.5, 2.0, 8.0])
    say("stack ftlm", lx, ly, nu, nd, "R", R, "M", M, "path", path, "blocks", len(blocks))
    for dtype in (L.FTLM_C128, L.FTLM_C64):
        ref, got = ritz_pair(blocks, R, M, dtype, path)
        assert np.array_equal(ref.used, got.used)
        scale = np.max(np.abs(ref.alpha))
        da = np.max(np.abs(ref.alpha - got.alpha)) / scale
        db = np.max(np.abs(ref.beta - got.beta)) / scale
        dl = np.max(np.abs(lnz_of(ref, betas) - lnz_of(got, betas)))
        say("  dtype", dtype, "d_alpha", da, "d_beta", db, "d_lnZ", dl)
        if dtype == L.FTLM_C128:
            assert da < 1e-10 and db < 1e-10 and dl < 1e-10
            assert np.max(np.abs(ref.ritz - got.ritz)) < 1e-9 * scale
        else:
            assert dl < 1e-5


def test_stack_ftlm_4x4_8_8_all_16_blocks():
    """All 16 momentum blocks of 4x4 (8,8) in one stack (R = 2 per block, M = 12, complex64 with
    FP64 reductions): alpha/beta and lnZ(beta) for beta in 0.1 .. 20 against the per-block path
    with the same starts (lnZ within 1e-6, alpha within 1e-5 relative: complex64 vectors summed
    in a different order)."""
    blocks = all_blocks(4, 4, 8, 8)
    betas = np.geomspace(0.1, 20.0, 8)
    ref, got = ritz_pair(blocks, 2, 12, L.FTLM_C64, L.FTLM_PATH_STACK, seed=17)
    assert np.array_equal(ref.used, got.used)
    scale = np.max(np.abs(ref.alpha))
    da = np.max(np.abs(ref.alpha - got.alpha)) / scale
    lz_ref = np.logaddexp.reduce(lnz_of(ref, betas), axis=0)
    lz_got = np.logaddexp.reduce(lnz_of(got, betas), axis=0)
    say("4x4 (8,8) stack ftlm: d_alpha", da, "block-summed lnZ", lz_got, "d", np.abs(lz_ref - lz_got).max())
    assert da < 1e-5
    assert np.max(np.abs(lz_ref - lz_got)) < 1e-6


def test_stack_krylov_session_matches_blocks():
    blocks = all_blocks(4, 2, 4, 4)
    nk, M = len(blocks), 20
    S = api.KStackOperator.from_blocks(blocks)
    ks = api.KrylovSession(S, nk * 2, M, dtype=L.FTLM_C128, start=L.FTLM_START_PHASE, seed=9)
    ks.run(M)
    a_s, b_s, u_s, _ = ks.result()
    for j, b in enumerate(blocks):
        kb = api.KrylovSession(b, 2, M, dtype=L.FTLM_C128, start=L.FTLM_START_PHASE, seed=9)
        kb.run(M)
        a_b, b_b, u_b, _ = kb.result()
        a_j, b_j = a_s[:, 2 * j:2 * j + 2], b_s[:, 2 * j:2 * j + 2]
        scale = np.max(np.abs(a_b))
        assert np.array_equal(u_b, u_s[2 * j:2 * j + 2])
        assert np.max(np.abs(a_j - a_b)) < 1e-10 * scale and np.max(np.abs(b_j - b_b)) < 1e-10 * scale


def test_stack_rejects_bad_inputs():
    a = all_blocks(4, 2, 4, 4)
    b = all_blocks(4, 2, 3, 4)  # another sector: another orbit graph
    with pytest.raises(ValueError):
        api.KStackOperator.from_blocks([a[0], b[0]])
    S = api.KStackOperator.from_blocks(a[:3])
    import torch

    x = torch.zeros(S.dim() * 4, dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError):
        S.apply_device(x, torch.empty_like(x), R=4, dtype=L.FTLM_C128)  # 4 is not a multiple of 3
    with pytest.raises(IndexError):
        S.scatter(3, x, x, rk=1)
