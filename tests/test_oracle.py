"""Pins the CPU oracle (oracle/ftlm_oracle.cpp) before it is trusted as the checker.

Every oracle function is checked against at least one of: the golden fixtures written by the
reference itself (tests/golden/make_golden.py), the reference's published data files, or what
the mathematics fixes (brute force, closed forms, numpy eigensolvers, identities).
"""

import itertools
import math

import numpy as np
import pytest

from pyoracle import params8


def model(golden, name):
    return params8(*golden[f"geom/{name}/model"])


# ------------------------------------------------------------------ geometry (src/lattice.cpp)

@pytest.mark.parametrize("name", ["c1_2x2", "pub_2x2", "flux_2x2", "c2_4x2", "flux_3x2", "c5_4x3", "c3_4x4"])
def test_geometry_bit_exact(golden, oracle, name):
    hops, pairs = oracle.geometry(model(golden, name))
    g = lambda k: golden[f"geom/{name}/{k}"]
    assert np.array_equal(hops["from"], g("hops_from"))
    assert np.array_equal(hops["to"], g("hops_to"))
    assert np.array_equal(hops["spin"], g("hops_spin"))
    assert np.array_equal(hops["re"] + 1j * hops["im"], g("hops_coeff"))
    assert np.array_equal(pairs, g("pairs"))


def test_geometry_duplicate_bonds_4x4(oracle):
    # Every diagonal bond is listed twice on Lx,Ly >= 3: summed NNN amplitude is -2t'.
    hops, pairs = oracle.geometry(params8(4, 4, 1.0, -0.25))
    assert len(hops) == 384 and len(pairs) == 32
    tot = {}
    for h in hops:
        k = (h["from"], h["to"], h["spin"])
        tot[k] = tot.get(k, 0) + h["re"]
    assert math.isclose(tot[(0, 1, 0)], -1.0) and math.isclose(tot[(0, 5, 0)], 0.5)


# ------------------------------------------------------------------ basis (src/fock_basis.cpp)

@pytest.mark.parametrize("key", ["4_2_2", "4_0_3", "8_4_4", "8_3_1", "12_2_1", "16_1_2", "6_3_3"])
def test_basis_and_index_of(golden, oracle, key):
    n, nu, nd = map(int, key.split("_"))
    up, dn = oracle.basis_masks(n, nu, nd)
    assert np.array_equal(up, golden[f"basis/{key}/up"])
    assert np.array_equal(dn, golden[f"basis/{key}/dn"])
    got = oracle.index_of(n, nu, nd, golden[f"basis/{key}/probe_up"], golden[f"basis/{key}/probe_dn"])
    assert np.array_equal(got, golden[f"basis/{key}/probe_idx"])


@pytest.mark.parametrize("n,k", [(4, 2), (6, 3), (9, 4), (12, 5), (16, 8), (16, 0), (16, 16), (5, 1)])
def test_basis_brute_force_and_closed_rank(oracle, n, k):
    # Lexicographic order on sorted position tuples == itertools.combinations order.
    want = [sum(1 << p for p in c) for c in itertools.combinations(range(n), k)]
    up, _ = oracle.basis_masks(n, k, 0)
    assert up.tolist() == want
    assert np.array_equal(oracle.rank_closed_form(n, k, up), np.arange(len(up)))


def test_basis_invalid_sector(oracle):
    assert oracle.basis_dim(0, 0, 0) == -1
    assert oracle.basis_dim(4, 5, 0) == -1
    assert oracle.basis_dim(33, 1, 1) == -1


# ------------------------------------------------------------------ H*v (src/hubbard_hamiltonian.cpp)

APPLY = [("c1_2x2", 2, 2), ("pub_2x2", 1, 2), ("flux_2x2", 2, 2), ("c2_4x2", 4, 4), ("c2_4x2", 3, 1),
         ("flux_3x2", 3, 2), ("c5_4x3", 2, 2), ("c3_4x4", 2, 1)]


@pytest.mark.parametrize("name,nu,nd", APPLY)
def test_apply_bit_exact(golden, oracle, name, nu, nd):
    key = f"apply/{name}/{nu}_{nd}"
    y = oracle.apply(model(golden, name), nu, nd, golden[key + "/x"])
    assert np.array_equal(y.view(np.float64), golden[key + "/y"].view(np.float64))


def test_apply_published_2x2_spectra(golden, oracle):
    # benchmarks/data/eigs_full_vs_k_2x2.tsv: sorted full-sector eigenvalues, all 25 sectors.
    rows = golden["published/eigs_2x2/rows"]
    p8 = model(golden, "pub_2x2")
    for nu in range(5):
        for nd in range(5):
            want = rows[(rows[:, 0] == nu) & (rows[:, 1] == nd)][:, 4]
            H = oracle.dense_hamiltonian(p8, nu, nd)
            assert np.abs(H - H.conj().T).max() < 1e-12
            got = np.linalg.eigvalsh(H)
            assert np.allclose(got, np.sort(want), atol=1e-9), (nu, nd)


def test_apply_closed_forms(oracle):
    # U only: H = U * double occupancy (diagonal);  4x4 single particle dispersion with the
    # doubled NNN list: eps(k) = -2t(cos kx + cos ky) - 8 t' cos kx cos ky.
    p8 = params8(2, 2, 0.0, 0.0, 3.0, 0.0)
    H = oracle.dense_hamiltonian(p8, 2, 1)
    up, dn = oracle.basis_masks(4, 2, 1)
    dbl = np.array([bin(int(u) & int(d)).count("1") for u, d in zip(up, dn)])
    assert np.allclose(H, np.diag(3.0 * dbl))
    t, tp = 1.0, -0.25
    H = oracle.dense_hamiltonian(params8(4, 4, t, tp), 1, 0)
    ks = [2 * math.pi * m / 4 for m in range(4)]
    eps = sorted(-2 * t * (math.cos(a) + math.cos(b)) - 8 * tp * math.cos(a) * math.cos(b) for a in ks for b in ks)
    assert np.allclose(np.linalg.eigvalsh(H), eps, atol=1e-12)


def test_apply_trace_identity(oracle):
    p8 = params8(4, 2, 1.0, -0.25, 8.0, 1.0)
    H = oracle.dense_hamiltonian(p8, 3, 2)
    up, dn = oracle.basis_masks(8, 3, 2)
    hops, pairs = oracle.geometry(p8)
    diag = []
    for u, d in zip(up.tolist(), dn.tolist()):
        occ = [((u >> s) & 1) + ((d >> s) & 1) for s in range(8)]
        diag.append(8.0 * bin(u & d).count("1") + 1.0 * sum(occ[i] * occ[j] for i, j in pairs))
    assert np.allclose(np.diag(H).real, diag)


# ------------------------------------------------------------------ Lanczos (src/lanczos.cpp)

@pytest.mark.parametrize("name,nu,nd", APPLY)
@pytest.mark.parametrize("seed", [7, 42])
def test_lanczos_bit_exact(golden, oracle, name, nu, nd, seed):
    key = f"lanczos/{name}/{nu}_{nd}/s{seed}"
    M = int(golden[key + "/M"][0])
    used, a, b = oracle.lanczos(model(golden, name), nu, nd, M, seed)
    assert used == int(golden[key + "/used"][0])
    assert np.array_equal(a, golden[key + "/alpha"])
    assert np.array_equal(b, golden[key + "/beta"])


def test_lanczos_exact_on_tiny_sector(oracle):
    # dim <= M: the Krylov space is the whole sector; capped Jacobi vs eigvalsh.
    p8 = params8(2, 2, 1.0, -0.2, 4.0, 0.5)
    used, a, b = oracle.lanczos(p8, 1, 0, 10, 3)
    assert used <= 4
    lam, _ = oracle.tridiag_eig_ql(a, b)
    H = oracle.dense_hamiltonian(p8, 1, 0)
    # a single start vector spans one copy of each distinct eigenvalue (degenerate pairs merge)
    distinct = np.unique(np.round(np.linalg.eigvalsh(H), 9))
    assert np.allclose(np.sort(lam), distinct, atol=1e-9)


def test_lanczos_supplied_start(oracle):
    p8 = params8(2, 2, 1.0, -0.2, 4.0, 0.5)
    q0 = oracle.gauss_start(36, 7)
    u1, a1, b1 = oracle.lanczos(p8, 2, 2, 20, 7)
    u2, a2, b2 = oracle.lanczos_from(p8, 2, 2, 20, q0)
    assert u1 == u2 and np.array_equal(a1, a2) and np.array_equal(b1, b2)


def test_extrema_config1(golden, oracle):
    used, lo, hi = oracle.lanczos_extrema(model(golden, "c1_2x2"), 2, 2, 36, 7)
    assert [used, lo, hi] == golden["extrema/c1_2x2/2_2/s7"].tolist()
    # converged values differ: the reference Jacobi stops after 120 rotations
    H = oracle.dense_hamiltonian(model(golden, "c1_2x2"), 2, 2).real
    ev = np.linalg.eigvalsh(H)
    assert abs(ev[0] - (-2.059486)) < 1e-6 and abs(ev[-1] - 16.870769) < 1e-6
    assert abs(lo - (-2.015624)) < 1e-6


def test_extrema_chain(golden, oracle):
    used, lo, hi = oracle.lanczos_extrema(params8(2, 1, 1.0), 1, 1, 48, 42)
    assert [used, lo, hi] == golden["extrema/chain_2x1/1_1/s42"].tolist()


def test_dense_extrema_1600(oracle):
    H = oracle.dense_hamiltonian(params8(2, 2, 1.0, -0.2, 4.0, 0.5), 2, 2).real
    lo, hi = oracle.dense_extrema(H, 1600)
    ev = np.linalg.eigvalsh(H)
    assert abs(lo - ev[0]) < 1e-6 and abs(hi - ev[-1]) < 1e-6


# ------------------------------------------------------------------ quadrature / FTLM (src/ftlm_thermo.cpp)

def test_ftlm_all_sectors_4x2(golden, oracle):
    p8 = model(golden, "c2_4x2")
    for real, key in ((False, "complex"), (True, "real")):
        want = golden[f"ftlm/c2_4x2/beta1_R2_M16_s7/{key}"]
        for nu in range(9):
            for nd in range(9):
                got = oracle.ftlm_log_partition(p8, nu, nd, 1.0, 2, 16, 7, real)
                assert got == want[nu, nd], (real, nu, nd)


@pytest.mark.parametrize("M", [80, 150])
@pytest.mark.parametrize("bt", [0.5, 2.0, 8.0])
def test_ftlm_capped_jacobi_long(golden, oracle, M, bt):
    got = oracle.ftlm_log_partition(model(golden, "c2_4x2"), 4, 4, bt, 1, M, 7)
    assert got == golden[f"ftlm/c2_4x2/4_4/M{M}_beta{bt}"][0]


def test_quadrature_beta0_is_log_norm(oracle):
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(30), np.abs(rng.standard_normal(30))
    assert abs(oracle.quadrature(a, b, 0.0)) < 1e-12  # sum of first-component weights is 1


def test_quadrature_matches_converged_when_converged(oracle):
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal(12), 0.5 + np.abs(rng.standard_normal(12))
    lam, w = oracle.tridiag_eig_ql(a, b)
    T = np.diag(a) + np.diag(b[:11], 1) + np.diag(b[:11], -1)
    ev, V = np.linalg.eigh(T)
    assert np.allclose(np.sort(lam), ev, atol=1e-12)
    assert np.allclose(np.sort(w), np.sort(V[0] ** 2), atol=1e-12)
    for bt in (0.3, 2.0):
        exact = np.log(np.sum(V[0] ** 2 * np.exp(-bt * ev)))
        assert abs(oracle.quadrature(a, b, bt) - exact) < 1e-10
        lq, wq = oracle.quadrature_ritz(a, b)
        assert abs(np.log(np.sum(wq * np.exp(-bt * lq))) - exact) < 1e-10


def test_ftlm_thermo_reference_test(golden):
    # tests/test_ftlm_thermo.cpp: diag(2,3), beta=1, R=48, M=32, seed 12345; tolerance 0.15.
    exact = math.log(math.exp(-2.0) + math.exp(-3.0))
    assert abs(golden["ftlm/diag23"][0] - exact) < 0.15


def test_gc_density_published_2x2(golden, oracle):
    # benchmarks/data/n_vs_mu_beta20.tsv (2x2 ED, beta=20): per-sector exact lnZ from dense
    # spectra, mixed with the bench's mu formula.
    p8 = model(golden, "pub_2x2")
    lnZ, N = [], []
    for nu in range(5):
        for nd in range(5):
            ev = np.linalg.eigvalsh(oracle.dense_hamiltonian(p8, nu, nd))
            m = (-20.0 * ev).max()
            lnZ.append(m + np.log(np.exp(-20.0 * ev - m).sum()))
            N.append(nu + nd)
    tab = golden["published/n_vs_mu_2x2_beta20"]
    got = oracle.gc_density(np.array(lnZ), np.array(N), 20.0, tab[:, 0], 4)
    assert np.allclose(got, tab[:, 1], rtol=2e-5, atol=1e-12)


def test_exact_log_partition(oracle):
    H = oracle.dense_hamiltonian(params8(2, 2, 1.0, -0.2, 4.0, 0.5, 0.3, -0.2), 2, 1)
    for bt in (0.5, 3.0):
        ev = np.linalg.eigvalsh(H)
        exact = np.log(np.exp(-bt * ev).sum())
        assert abs(oracle.exact_log_partition(H, bt) - exact) < 1e-9


# ------------------------------------------------------------------ orbit k-blocks

@pytest.mark.parametrize("lx,ly,nu,nd", [(2, 2, 2, 2), (3, 2, 3, 3), (4, 2, 4, 4), (4, 2, 2, 3)])
def test_orbit_basis_bit_exact(golden, oracle, lx, ly, nu, nd):
    dims = golden[f"orbit/{lx}x{ly}/{nu}_{nd}/dims"]
    for ky in range(ly):
        for kx in range(lx):
            reps, osz, ssz, A = oracle.orbit_basis(lx, ly, nu, nd, kx, ky)
            assert len(reps) == dims[ky, kx]
            key = f"orbit/{lx}x{ly}/{nu}_{nd}/k{kx}{ky}"
            if key + "/reps" in golden:
                assert np.array_equal(reps, golden[key + "/reps"])
                assert np.array_equal(osz, golden[key + "/orbit_size"])
                assert np.array_equal(ssz, golden[key + "/stab_size"])
                assert np.array_equal(A, golden[key + "/A"])
    # Sigma_K dim_K = dim_full; orbit |orbit|*|stab| = |G|
    assert dims.sum() == oracle.basis_dim(lx * ly, nu, nd)


def test_orbit_reference_test_values(oracle):
    # tests/test_orbit_k_block_basis.cpp prints dims 12 (2x2 (2,2) Gamma) and 66 (3x2 (3,3) K=(1,0)).
    assert len(oracle.orbit_basis(2, 2, 2, 2, 0, 0)[0]) == 12
    assert len(oracle.orbit_basis(3, 2, 3, 3, 1, 0)[0]) == 66
    reps, osz, ssz, A = oracle.orbit_basis(2, 2, 2, 2, 0, 0)
    assert np.all(osz * ssz == 4)
    assert np.allclose(np.abs(A), ssz)


@pytest.mark.parametrize("case", ["3x2/3_3/k10", "2x2/2_2/k00", "4x2/4_4/k21", "4x2/2_3/k10"])
def test_orbit_apply_and_map_bit_exact(golden, oracle, case):
    shape, sec, k = case.split("/")
    lx, ly = map(int, shape.split("x"))
    nu, nd = map(int, sec.split("_"))
    kx, ky = int(k[1]), int(k[2])
    key = f"orbit_apply/{case}"
    p8 = params8(lx, ly, 1.0, -0.25, 8.0, 1.0)
    y = oracle.orbit_apply(p8, nu, nd, kx, ky, golden[key + "/x"])
    assert np.array_equal(y.view(np.float64), golden[key + "/y"].view(np.float64))
    up, dn = oracle.basis_masks(lx * ly, nu, nd)
    st = ((up.astype(np.uint32) << np.uint32(16)) | dn.astype(np.uint32)).astype(np.uint32)
    row, dx, dy, eta = oracle.orbit_map(lx, ly, nu, nd, kx, ky, st)
    assert np.array_equal(np.stack([row, dx, dy, eta.astype(np.int32)], axis=1), golden[key + "/map"])
    lz = oracle.orbit_ftlm_log_partition(p8, nu, nd, kx, ky, 1.5, 2, 12, 3)
    assert lz == golden[key + "/lnZ_b1.5_R2_M12_s3"][0]


def test_orbit_blocks_reproduce_full_spectrum(oracle):
    # union over K of the orbit-block spectra == full-sector spectrum (phi = 0)
    lx, ly, nu, nd = 3, 2, 2, 2
    p8 = params8(lx, ly, 1.0, -0.3, 5.0, 0.7)
    full = np.linalg.eigvalsh(oracle.dense_hamiltonian(p8, nu, nd))
    evs = []
    for ky in range(ly):
        for kx in range(lx):
            dk = len(oracle.orbit_basis(lx, ly, nu, nd, kx, ky)[0])
            Hk = np.stack([oracle.orbit_apply(p8, nu, nd, kx, ky, e) for e in np.eye(dk, dtype=np.complex128)], 1)
            assert np.abs(Hk - Hk.conj().T).max() < 1e-12
            evs.extend(np.linalg.eigvalsh(Hk))
    assert np.allclose(np.sort(evs), full, atol=1e-10)


def test_translation_sign_closed_form(oracle):
    # eta for a translation of a single spin species equals the parity of the cyclic shift of
    # the occupied modes: brute-force the permutation parity independently.
    rng = np.random.default_rng(11)
    for _ in range(200):
        u, d = int(rng.integers(0, 1 << 8)), int(rng.integers(0, 1 << 8))
        dx, dy = int(rng.integers(0, 4)), int(rng.integers(0, 2))
        modes = []
        for s in range(8):
            t = ((s // 4 + dy) % 2) * 4 + (s % 4 + dx) % 4
            if (u >> s) & 1:
                modes.append(2 * t)
            if (d >> s) & 1:
                modes.append(2 * t + 1)
        perm = np.argsort(modes)
        visited, parity = [False] * len(perm), 0
        for i in range(len(perm)):
            if not visited[i]:
                j, c = i, 0
                while not visited[j]:
                    visited[j] = True
                    j = perm[j]
                    c += 1
                parity += c - 1
        assert oracle.translation_sign(u, d, 4, 2, dx, dy) == (-1.0 if parity % 2 else 1.0)


def test_apply_rows_equals_full_push(oracle):
    # the row-restricted push used for 4x4 (8,8) checks reproduces the full push bit for bit
    p8 = params8(4, 3, 1.0, -0.25, 8.0, 1.0, 0.1, 0.0)
    d = oracle.basis_dim(12, 3, 2)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    y = oracle.apply(p8, 3, 2, x)
    rows = rng.integers(0, d, 300)
    got = oracle.apply_rows(p8, 3, 2, x, rows)
    assert np.array_equal(got.view(np.float64), y[rows].view(np.float64))


def test_oracle_translate_group_action(oracle):
    """orc_translate is a Z_lx x Z_ly action: composition adds shifts, full cycles are the
    identity, occupation counts are kept, and the canonical state is the orbit's minimum."""
    rng = np.random.default_rng(3)
    for (lx, ly) in [(4, 4), (4, 2), (3, 2)]:
        n = lx * ly
        for _ in range(40):
            u = int(rng.integers(0, 1 << n))
            d = int(rng.integers(0, 1 << n))
            base = (u << 16) | d
            assert oracle.translate(u, d, lx, ly, lx, 0) == base
            assert oracle.translate(u, d, lx, ly, 0, ly) == base
            imgs = []
            for dx in range(lx):
                for dy in range(ly):
                    t = oracle.translate(u, d, lx, ly, dx, dy)
                    assert bin(t >> 16).count("1") == bin(u).count("1")
                    assert bin(t & 0xFFFF).count("1") == bin(d).count("1")
                    t2 = oracle.translate(t >> 16, t & 0xFFFF, lx, ly, 1, 2)
                    assert t2 == oracle.translate(u, d, lx, ly, dx + 1, dy + 2)
                    imgs.append(t)
            assert oracle.canonical(u, d, lx, ly) == min(imgs)


def _golden_lz_close(got, want):
    """The reference printed logZ with 6 significant digits (and -0 / -inf)."""
    if not math.isfinite(want):
        return not math.isfinite(got)
    return abs(got - want) <= 6e-6 * max(abs(want), 1.0)


def test_kgram_blocks_reproduce_reference_run(oracle):
    """The oracle's restatement of the reference's Gram momentum blocks (kgram_block) against the
    reference's own 3x2 run (benchmarks/data/n_vs_mu_3x2_beta20_four_solver_monitoring.txt):
    per sector, the summed block dims (orbits with an odd fermionic stabiliser sign drop out, e.g.
    12 of 15 states in (0,2), none in (0,6)), and ln sum_K tr e^{-beta H_K} by the reference's
    capped-Jacobi ED against the printed logZ of every sector the k-gram run solved by ED and of
    every sector of the ed run (Gram blocks, all by ED)."""
    import json
    import os

    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "nmu3x2_solvers.json")))
    p8 = params8(3, 2, 1.0, -0.35, 5.75, 0.9)
    kg, ed = d["solvers"]["k-gram"]["sectors"], d["solvers"]["ed"]["sectors"]
    dropped = 0
    for key, (dim_full, dk_tot, n_ed, n_ftlm, lz) in kg.items():
        nu, nd = map(int, key.split(","))
        blocks = [oracle.kgram_block(p8, nu, nd, kx, ky) for ky in range(2) for kx in range(3)]
        dims = [b.shape[0] for b in blocks]
        assert sum(dims) == dk_tot == ed[key][1], key
        assert max(dims) <= 64 or n_ftlm > 0
        dropped += dim_full - dk_tot
        lzs = [oracle.exact_log_partition(b, 20.0) for b in blocks if b.shape[0] > 0]
        ours = -math.inf
        for v in lzs:
            ours = float(np.logaddexp(ours, v))
        assert _golden_lz_close(ours, ed[key][4]), (key, ours, ed[key][4])
        if n_ftlm == 0:
            assert _golden_lz_close(ours, lz), (key, ours, lz)
    assert dropped == 3 * 4 + 9 * 2 + 1 * 2  # (0,2),(2,0),(4,6),(6,4); (2,4),(4,2); (0,6),(6,0)


@pytest.mark.parametrize("lx,ly", [(2, 2), (3, 2)])
def test_kgram_dims_against_dense_translation_projector(oracle, lx, ly):
    """Mirrors tests/test_k_gram_mismatch_scan.cpp: the Gram block dimension of every sector and
    momentum against rank(P_K) of the dense translation projector P_K = (1/N) sum_t e^{-iK.t} T(t).
    What the mathematics fixes and the projector does not take from the Gram code: T(t) is a
    unitary representation of Z_lx x Z_ly that commutes with H; P_K are orthogonal projectors
    that sum to the identity, so sum_K rank(P_K) = dim; the Gram subspace lies inside range(P_K),
    so k_out <= rank(P_K), and where they are equal the Gram block has the spectrum of H on
    range(P_K).  Where they differ are the orbits whose bosonic character and fermionic sign are
    both non-trivial on the stabiliser (the product is trivial): the reference's seed rule leaves
    them out.  On 3x2 these are the sectors and counts its own 3x2 run shows (3 of 15 states in
    (0,2), the single state of (0,6), and their particle-hole / spin images; (2,4) and (4,2) are
    past the scan's dim <= 64); on 2x2 every state of (0,2), (2,0), (2,4) and (4,2) is left out,
    since each two-fermion configuration there is swapped onto itself by one translation."""
    import symmetry_kat as kat

    n = lx * ly
    p8 = params8(lx, ly, 1.0, -0.35, 5.75, 0.9)
    shifts = [(dx, dy) for dy in range(ly) for dx in range(lx)]
    missing = {}
    for nu in range(n + 1):
        for nd in range(n + 1):
            up, dn = oracle.basis_masks(n, nu, nd)
            dim = len(up)
            if dim == 0 or dim > 64:
                continue
            H = oracle.dense_hamiltonian(p8, nu, nd)
            T = {s: kat.dense_translation(oracle, lx, ly, up, dn, *s) for s in shifts}
            for (ax, ay), Ta in T.items():
                assert np.array_equal(Ta.T @ Ta, np.eye(dim))
                assert np.abs(Ta @ H - H @ Ta).max() < 1e-12, (nu, nd, ax, ay)
                for (bx, by), Tb in T.items():
                    assert np.array_equal(Ta @ Tb, T[(ax + bx) % lx, (ay + by) % ly])
            full = np.linalg.eigvalsh(H)
            total, spectra, psum = 0, [], np.zeros((dim, dim), complex)
            for ky in range(ly):
                for kx in range(lx):
                    P = sum(np.exp(-2j * math.pi * (kx * dx / lx + ky * dy / ly)) * T[dx, dy]
                            for dx, dy in shifts) / n
                    psum += P
                    assert np.abs(P @ P - P).max() < 1e-12 and np.abs(P - P.conj().T).max() < 1e-12
                    w, Q = np.linalg.eigh(P)
                    Q = Q[:, w > 0.5]
                    rank = Q.shape[1]
                    total += rank
                    spectra.extend(np.linalg.eigvalsh(Q.conj().T @ H @ Q))
                    B = oracle.kgram_block(p8, nu, nd, kx, ky)
                    k_out = B.shape[0]
                    assert k_out <= rank, (nu, nd, kx, ky)
                    if rank > k_out:
                        missing[nu, nd] = missing.get((nu, nd), 0) + rank - k_out
                    if k_out == rank and rank > 0:
                        assert np.allclose(np.linalg.eigvalsh(B), np.linalg.eigvalsh(Q.conj().T @ H @ Q),
                                           atol=1e-10), (nu, nd, kx, ky)
            assert np.abs(psum - np.eye(dim)).max() < 1e-12
            assert total == dim
            assert np.allclose(np.sort(spectra), full, atol=1e-10)
    want = {(2, 2): {(0, 2): 6, (2, 0): 6, (2, 4): 6, (4, 2): 6},
            (3, 2): {(0, 2): 3, (2, 0): 3, (4, 6): 3, (6, 4): 3, (0, 6): 1, (6, 0): 1}}[lx, ly]
    assert missing == want, missing


def test_symmetry_known_answers(oracle):
    """The reference's translation-symmetry known answers (test_symmetry_translation_stage1.cpp,
    test_symmetry_stage2.cpp, test_symmetry_stage3.cpp, test_momentum_partition_2x2.cpp) on the
    oracle's translate / sign / canonical routines."""
    import cmath

    import symmetry_kat as kat

    def translate(lx, ly, u, d, dx, dy):
        return int(oracle.translate(u, d, lx, ly, dx, dy)), oracle.translation_sign(u, d, lx, ly, dx, dy)

    def canonical(lx, ly, u, d):
        c = int(oracle.canonical(u, d, lx, ly))
        return c, sum(int(oracle.translate(u, d, lx, ly, dx, dy)) == ((u << 16) | d)
                      for dy in range(ly) for dx in range(lx))

    def bloch(lx, ly, kx, ky, dx, dy):
        return cmath.exp(-2j * math.pi * (kx * dx / lx + ky * dy / ly))

    kat.check_translation_stage1(translate)
    kat.check_symmetry_stage2(translate, canonical)
    kat.check_symmetry_stage3(translate, canonical, bloch)
    kat.check_momentum_partition_2x2(translate, canonical, oracle.basis_masks)


@pytest.mark.parametrize("lx,ly,nu,nd,kx,ky,phi", [(4, 2, 4, 4, 1, 0, 0.0), (3, 2, 3, 3, 1, 0, 0.0),
                                                   (4, 2, 2, 3, 1, 1, 0.3), (4, 3, 3, 2, 2, 1, 0.0)])
def test_orbit_apply_rows_equals_block_push(oracle, lx, ly, nu, nd, kx, ky, phi):
    """The row-restricted orbit apply (each row's pushes enumerated from its own orbit, no sector
    map) reproduces the oracle's full block push (itself bit-exact against the reference sources),
    flux included; a representative list that disagrees with the basis is refused."""
    p8 = params8(lx, ly, 1.0, -0.25, 8.0, 1.0, phi, 0.0)
    reps, _, _, _ = oracle.orbit_basis(lx, ly, nu, nd, kx, ky)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(len(reps)) + 1j * rng.standard_normal(len(reps))
    y = oracle.orbit_apply(p8, nu, nd, kx, ky, x)
    rows = rng.choice(len(reps), size=min(len(reps), 200), replace=False)
    yr = oracle.orbit_apply_rows(p8, nu, nd, kx, ky, reps, x, rows)
    assert np.max(np.abs(yr - y[rows])) < 1e-13 * np.max(np.abs(y))
    with pytest.raises(AssertionError, match="disagree"):
        oracle.orbit_apply_rows(p8, nu, nd, kx, ky, reps[1:], x[1:], np.arange(len(reps) - 1))


@pytest.mark.parametrize("lx,ly,nu,nd,kx,ky", [(4, 2, 4, 4, 1, 0), (3, 2, 3, 3, 1, 0), (4, 2, 2, 3, 1, 1),
                                               (2, 2, 2, 2, 0, 0), (4, 3, 3, 2, 2, 1)])
def test_orbit_lift_is_an_isometry_intertwining_h(oracle, lx, ly, nu, nd, kx, ky):
    """Phi_K (columns: the reference's Bloch sums of each representative) is an isometry and
    H Phi_K v = Phi_K H_K v at zero flux, with H the sector push and H_K the block push.  This is
    the identity behind the lifted-start Lanczos check of the momentum blocks; a wrong phase
    convention, sign or normalisation breaks it."""
    p8 = params8(lx, ly, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0)
    reps, _, _, _ = oracle.orbit_basis(lx, ly, nu, nd, kx, ky)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(len(reps)) + 1j * rng.standard_normal(len(reps))
    f = oracle.orbit_lift(lx, ly, nu, nd, kx, ky, reps, v)
    assert abs(np.linalg.norm(f) - np.linalg.norm(v)) < 1e-13 * np.linalg.norm(v)
    hf = oracle.apply(p8, nu, nd, f)
    fh = oracle.orbit_lift(lx, ly, nu, nd, kx, ky, reps, oracle.orbit_apply(p8, nu, nd, kx, ky, v))
    assert np.max(np.abs(hf - fh)) < 1e-13 * np.max(np.abs(hf))
    # a conjugated phase convention does not intertwine (except where every phase is real)
    if (2 * kx) % lx or (2 * ky) % ly:
        g = oracle.orbit_lift(lx, ly, nu, nd, (-kx) % lx, (-ky) % ly, reps, v)
        assert np.max(np.abs(oracle.apply(p8, nu, nd, g) - fh)) > 1e-3
