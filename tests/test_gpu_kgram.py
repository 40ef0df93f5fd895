"""GPU: momentum Gram blocks (HubbardMomentumBlock, ftlm_hip_kgram_create) and the k-gram / k-orbit
/ ed solvers of ftlm_nmu_rect_k against the reference's own 3x2 run.

- Blocks: the device block's dimension equals the oracle's Gram block (kgram_block), its dense
  matrix (applied to unit vectors) is Hermitian to 1e-12 and its spectrum equals the oracle block's
  to 1e-10 (both are Phi^dag H Phi in orthonormal bases of one subspace, so only spectra compare).
- Drivers: tests/golden/nmu3x2_solvers.json holds what the reference's bench_ftlm_nmu_rect_k
  printed for --solver=k-gram, k-orbit and ed (3x2, beta=20, R=12, M=72, seed 7).  Every sector's
  dim_k_total must be equal, and logZ equal to the printed 6 digits (6e-6 relative) wherever the
  reference solved every block by ED.  FTLM blocks of the Gram solver draw their starts in the
  reference's zheev basis, which no other LAPACK build reproduces: the (3,3) sector is checked
  against the exact Gram-block value within the FTLM spread seen in the reference's own runs, and
  n(mu) within 0.05 of the reference's table (its four solvers differ by up to 0.019).
"""
from __future__ import annotations

import json
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from pyoracle import params8  # noqa: E402

pytestmark = pytest.mark.gpu
api = pytest.importorskip("ftlm_extended_hubbard_amd.api")
from ftlm_extended_hubbard_amd import _lib as L  # noqa: E402

BIN = os.path.join(ROOT, "ftlm_extended_hubbard_amd", "bin")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "nmu3x2_solvers.json")))
SECTOR_RE = re.compile(r"^sector \((\d+),(\d+)\) dim_full=(\d+) dim_k_total=(\d+) ed_blocks=(\d+) "
                       r"ftlm_blocks=(\d+) logZ=(\S+)$")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if api.device_count() < 1:
        pytest.skip("no HIP device")


def dense(op):
    d = op.dim()
    H = np.zeros((d, d), complex)
    for j in range(d):
        e = np.zeros(d, complex)
        e[j] = 1.0
        H[:, j] = op.apply(e)
    return H


@pytest.mark.parametrize("lx,ly,nu,nd,phi", [(3, 2, 2, 4, 0.0), (3, 2, 3, 3, 0.0), (3, 2, 0, 2, 0.0),
                                             (4, 2, 2, 2, 0.0), (4, 2, 3, 1, 0.3)])
def test_gram_blocks_match_oracle(oracle, lx, ly, nu, nd, phi):
    p = api.HubbardParams(lx, ly, 1.0, -0.35, 5.75, 0.9, phi, 0.0)
    p8 = params8(lx, ly, 1.0, -0.35, 5.75, 0.9, phi, 0.0)
    sector = api.HubbardOperator.from_params(p, nu, nd)
    first = None
    total = 0
    for ky in range(ly):
        for kx in range(lx):
            say("gram block", lx, ly, nu, nd, "K", kx, ky, "phi", phi)
            K = api.MomentumSector(kx, ky, lx, ly)
            kb = api.KBlockOperator.from_params(p, nu, nd, K, share=first)
            first = first or kb
            g = api.KGramOperator.from_params(p, nu, nd, K, sector=sector, share=first)
            want = oracle.kgram_block(p8, nu, nd, kx, ky)
            assert g.dim() == want.shape[0]
            total += g.dim()
            if g.dim() == 0:
                continue
            H = dense(g)
            scale = max(1.0, np.max(np.abs(H)))
            assert np.max(np.abs(H - H.conj().T)) < 1e-12 * scale
            assert np.max(np.abs(np.linalg.eigvalsh(H) - np.linalg.eigvalsh(want))) < 1e-10 * scale
    assert total <= sector.dim()


def test_gram_block_block_apply_and_c64(oracle):
    """Blocks of R vectors equal R single applies (complex64 storage to single precision), and the
    device ED of a block larger than one 32-vector round of the sector apply agrees with the exact
    ln tr e^{-beta H_K} of the oracle's block."""
    import torch

    p = api.HubbardParams(4, 2, 1.0, -0.35, 5.75, 0.9, 0.0, 0.0)
    sector = api.HubbardOperator.from_params(p, 3, 3)
    K = api.MomentumSector(1, 1, 4, 2)
    g = api.KGramOperator.from_params(p, 3, 3, K, sector=sector)
    d = g.dim()
    say("gram block apply R=32/8, c128/c64, 4x2 (3,3) K=(1,1), dim", d)
    rng = np.random.default_rng(3)
    for R in (32, 8):
        x = rng.standard_normal((d, R)) + 1j * rng.standard_normal((d, R))
        ref = np.stack([g.apply(x[:, r]) for r in range(R)], axis=1)
        for dt, tol in ((L.FTLM_C128, 1e-13), (L.FTLM_C64, 2e-6)):
            ct = np.complex64 if dt == L.FTLM_C64 else np.complex128
            xt = torch.from_numpy(x.astype(ct).reshape(-1)).to("cuda")
            yt = torch.empty_like(xt)
            g.apply_device(xt, yt, R=R, dtype=dt)
            torch.cuda.synchronize()
            y = yt.cpu().numpy().reshape(d, R)
            assert np.max(np.abs(y - ref)) <= tol * np.max(np.abs(ref)), (R, dt)
    # device ED (the reference's capped Jacobi, max(5000, 150 m) rotations) of a 3x2 (2,3) block
    # of 50 states: two rounds of the sector apply; the cap leaves ~1e-6 relative in lnZ
    p3 = api.HubbardParams(3, 2, 1.0, -0.35, 5.75, 0.9, 0.0, 0.0)
    g3 = api.KGramOperator.from_params(p3, 2, 3, api.MomentumSector(0, 0, 3, 2))
    w = np.linalg.eigvalsh(oracle.kgram_block(params8(3, 2, 1.0, -0.35, 5.75, 0.9), 2, 3, 0, 0))
    assert len(w) == g3.dim() and g3.dim() > 32
    betas = np.array([0.5, 2.0])
    got = api.exact_log_partition(g3, betas)
    for b, v in zip(betas, got):
        mx = np.max(-b * w)
        assert abs(v - (mx + np.log(np.sum(np.exp(-b * w - mx))))) < 1e-5 * max(1.0, abs(v))


def run_driver(solver, tmp_path):
    out = tmp_path / f"{solver}.tsv"
    args = [os.path.join(BIN, "ftlm_nmu_rect_k"), *GOLD["args"], f"--solver={solver}", "--no-monitor=1", f"--out={out}"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    sectors = {}
    for line in r.stderr.splitlines():
        m = SECTOR_RE.match(line.strip())
        if m:
            sectors[f"{m.group(1)},{m.group(2)}"] = [int(m.group(3)), int(m.group(4)), int(m.group(5)),
                                                     int(m.group(6)), float(m.group(7))]
    rows = np.array([[float(v) for v in line.split()] for line in out.read_text().splitlines()
                     if line.strip() and not line.startswith("#")])
    return sectors, rows, r.stdout


def lz_close(got, want):
    if not math.isfinite(want):
        return not math.isfinite(got)
    return abs(got - want) <= 6e-6 * max(abs(want), 1.0)


@pytest.mark.parametrize("solver", ["k-gram", "k-orbit", "ed"])
def test_driver_matches_reference_run(solver, tmp_path):
    say("ftlm_nmu_rect_k --solver=" + solver, *GOLD["args"])
    gold = GOLD["solvers"][solver]
    sectors, rows, stdout = run_driver(solver, tmp_path)
    assert f"METRIC kind=FTLM_nmu_rect_k solver={solver}" in stdout
    assert sectors.keys() == gold["sectors"].keys()
    for key, (df, dk, ne, nf, lz) in gold["sectors"].items():
        got = sectors[key]
        assert got[:4] == [df, dk, ne, nf], (key, got, gold["sectors"][key])
        if nf == 0:
            assert lz_close(got[4], lz), (key, got[4], lz)
    # the FTLM sector (3,3): within the spread of the reference's own FTLM runs of it around the
    # exact Gram-block value (k-gram -33.5301, k-orbit -34.0336, exact -33.8902)
    exact33 = GOLD["solvers"]["ed"]["sectors"]["3,3"][4]
    assert abs(sectors["3,3"][4] - exact33) < 1.0, sectors["3,3"]
    ref_rows = np.array(gold["n_mu"])
    assert np.allclose(rows[:, 0], ref_rows[:, 0], atol=1e-4)
    assert np.max(np.abs(rows[:, 1] - ref_rows[:, 1])) < 0.05


def test_driver_default_solver_is_k_gram(tmp_path):
    args = [os.path.join(BIN, "ftlm_nmu_rect_k"), "--Lx=2", "--Ly=2", "--beta=2", "--n-mu=5", "--no-monitor=1",
            f"--out={tmp_path / 'd.tsv'}"]
    say("ftlm_nmu_rect_k default solver 2x2")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "METRIC kind=FTLM_nmu_rect_k solver=default" in r.stdout and "[solver]" not in r.stdout
    r2 = subprocess.run([*args[:-1], f"--out={tmp_path / 'g.tsv'}", "--solver=k-gram"], capture_output=True,
                        text=True, timeout=120)
    assert r2.returncode == 0
    assert (tmp_path / "d.tsv").read_text().splitlines()[2:] == (tmp_path / "g.tsv").read_text().splitlines()[2:]


def test_orbit_all_keeps_every_state(oracle, tmp_path):
    """--solver=k-orbit --orbit-all=1 (beyond the reference): orbit blocks in every sector, so
    every sector's blocks cover all its states and the all-ED sectors give ln tr e^{-beta H}."""
    args = [os.path.join(BIN, "ftlm_nmu_rect_k"), "--Lx=3", "--Ly=2", "--beta=1.5", "--n-mu=11", "--no-monitor=1",
            "--solver=k-orbit", "--orbit-all=1", f"--out={tmp_path / 'o.tsv'}"]
    say("ftlm_nmu_rect_k --solver=k-orbit --orbit-all=1 3x2")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    p8 = params8(3, 2, 1.0, -0.35, 5.75, 0.9)
    seen = 0
    for line in r.stderr.splitlines():
        m = SECTOR_RE.match(line.strip())
        if not m:
            continue
        nu, nd, df, dk, ne, nf, lz = m.groups()
        assert int(df) == int(dk), line
        seen += 1
        if int(nf) == 0:
            w = np.linalg.eigvalsh(oracle.dense_hamiltonian(p8, int(nu), int(nd)))
            mx = np.max(-1.5 * w)
            exact = mx + np.log(np.sum(np.exp(-1.5 * w - mx)))
            assert abs(float(lz) - exact) <= 1e-5 * max(1.0, abs(exact)), line
    assert seen == 49


# ---- the reference's Gram-path tests (they need LAPACK there), on the device Gram blocks

def _dense_from(op):
    d = op.dim()
    return np.stack([op.apply(np.eye(d, dtype=complex)[:, j]) for j in range(d)], axis=1) if d else np.zeros((0, 0))


def test_single_particle_dispersion_4x4():
    """tests/test_hubbard_momentum_action.cpp:29-66 and test_lanczos_engine.cpp:34-52: one up
    electron on 4x4 (t=1, t'=U=V=0): every momentum block has one state with energy
    -2t(cos(pi kx/2) + cos(pi ky/2)) (1e-10), and 24-step Lanczos extrema equal it (1e-8), for the
    Gram and the orbit blocks."""
    p = api.HubbardParams(4, 4, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    sector = api.HubbardOperator.from_params(p, 1, 0)
    say("single-particle dispersion 4x4")
    for ky in range(4):
        for kx in range(4):
            K = api.MomentumSector(kx, ky, 4, 4)
            eps = -2.0 * (math.cos(0.5 * math.pi * kx) + math.cos(0.5 * math.pi * ky))
            for op in (api.KGramOperator.from_params(p, 1, 0, K, sector=sector),
                       api.KBlockOperator.from_params(p, 1, 0, K)):
                assert op.dim() == 1
                y = op.apply(np.array([1.0 + 0j]))
                assert abs(y[0].real - eps) < 1e-10 and abs(y[0].imag) < 1e-10
                ex = api.lanczos_extrema(op, 24, 3 + kx + 4 * ky)
                assert abs(ex.min_eval - eps) < 1e-8 and abs(ex.max_eval - eps) < 1e-8


def test_two_level_gamma_block():
    """tests/test_lanczos_engine.cpp:54-70: t=0, U=1, one up and one down electron on 4x4: the
    Gamma block holds the doubly occupied orbit (energy U) and the others (energy 0), so its
    Lanczos extrema are 0 and 1 (1e-8)."""
    p = api.HubbardParams(4, 4, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)
    say("two-level Gamma block 4x4 (1,1)")
    g = api.KGramOperator.from_params(p, 1, 1, api.MomentumSector(0, 0, 4, 4))
    ex = api.lanczos_extrema(g, 32, 99)
    assert abs(ex.min_eval - 0.0) < 1e-8 and abs(ex.max_eval - 1.0) < 1e-8


@pytest.mark.parametrize("lx,ly,tp,U,V,nu,nd,kx,ky", [
    (2, 2, -0.2, 4.0, 0.5, 1, 1, 0, 0), (2, 2, -0.2, 4.0, 0.5, 1, 1, 1, 0), (2, 2, -0.2, 4.0, 0.5, 2, 1, 1, 1),
    (3, 2, -0.35, 5.75, 0.9, 2, 3, 1, 0)])
def test_orbit_block_equals_gram_block(lx, ly, tp, U, V, nu, nd, kx, ky):
    """tests/test_orbit_k_block_apply_vs_production.cpp:135-207: the orbit block and the Gram block
    are one operator in two bases (the reference compares them in the full space); here their
    dimensions and spectra agree (1e-10)."""
    p = api.HubbardParams(lx, ly, 1.0, tp, U, V, 0.0, 0.0)
    K = api.MomentumSector(kx, ky, lx, ly)
    say("orbit vs gram", lx, ly, nu, nd, kx, ky)
    o = api.KBlockOperator.from_params(p, nu, nd, K)
    g = api.KGramOperator.from_params(p, nu, nd, K)
    assert o.dim() == g.dim() > 0
    wo = np.sort(np.linalg.eigvals(_dense_from(o)).real)
    wg = np.linalg.eigvalsh(_dense_from(g))
    assert np.max(np.abs(wo - wg)) < 1e-10 * max(1.0, np.max(np.abs(wg)))


def test_momentum_trace_2x2():
    """tests/test_momentum_trace_2x2.cpp:77-125: on 2x2 (t=1, t'=-0.35, U=5.75, V=0.9), for every
    sector with a momentum block, the Gram blocks' traces sum to tr H of the sector (1e-8)."""
    p = api.HubbardParams(2, 2, 1.0, -0.35, 5.75, 0.9, 0.0, 0.0)
    say("momentum trace 2x2")
    checked = 0
    for nu in range(5):
        for nd in range(5):
            sector = api.HubbardOperator.from_params(p, nu, nd)
            blocks = [api.KGramOperator.from_params(p, nu, nd, api.MomentumSector(kx, ky, 2, 2), sector=sector)
                      for ky in range(2) for kx in range(2)]
            if sum(b.dim() for b in blocks) == 0:
                continue
            tr_k = sum(np.trace(_dense_from(b)).real for b in blocks if b.dim())
            tr = np.trace(_dense_from(sector)).real
            assert abs(tr - tr_k) < 1e-8, (nu, nd, tr, tr_k)
            checked += 1
    assert checked >= 20


# The readers of the reference's plot scripts, restated: the TSV loader and its header parser
# (benchmarks/plot_n_vs_mu.py:27-52, also used by plot_n_vs_mu_3x2_four_solver.py:112-118) and the
# METRIC patterns (plot_n_vs_mu_3x2_four_solver.py:28-33, plot_n_vs_mu_3x2_three_way.py:21-27).
_HDR_NUM = r"=([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)"
_METRIC_FOUR = re.compile(
    r"^METRIC\s+kind=\S+\s+solver=(?P<solver>[^\s]+)\s+"
    r".*?wall_time_s=(?P<wall>[\d.eE+-]+)\s+.*?wall_sector_logZ_s=(?P<wall_sec>[\d.eE+-]+)\s+"
    r".*?peak_rss_bytes=(?P<rss_b>[\d.eE+-]+)\s+.*?peak_rss_mib=(?P<rss_mib>[\d.eE+-]+)", re.DOTALL)
_METRIC_THREE = re.compile(
    r"^METRIC\s+kind=(?P<kind>\S+)\s+.*?wall_time_s=(?P<wall>[\d.eE+-]+)\s+.*?"
    r"peak_rss_bytes=(?P<rss_b>[\d.eE+-]+)\s+.*?peak_rss_mib=(?P<rss_mib>[\d.eE+-]+)", re.DOTALL)


def _plot_load_tsv(text):
    mus, ns, meta = [], [], {}
    for line in text.splitlines():
        line = line.strip()
        if not line:
            continue
        if line.startswith("#"):
            for key in ("beta", "t", "tp", "U", "V"):
                m = re.search(rf"\b{key}" + _HDR_NUM, line)
                if m:
                    meta[key] = float(m.group(1))
            continue
        parts = line.split()
        mus.append(float(parts[0]))
        ns.append(float(parts[1]))
    return np.array(mus), np.array(ns), meta


@pytest.mark.parametrize("driver", ["ftlm_nmu_rect", "ftlm_nmu_rect_k"])
def test_driver_output_reads_with_reference_plot_scripts(driver, tmp_path):
    """The TSV and METRIC line of both drivers (monitoring on, as the reference's plot workflows run
    them) parse with the reference's plot-script readers into the run's model, grid and timings."""
    out = tmp_path / "n.tsv"
    args = [a for a in GOLD["args"] if driver == "ftlm_nmu_rect_k" or not a.startswith("--no-log-k-dims")]
    if driver == "ftlm_nmu_rect_k":
        args.append("--solver=k-orbit")
    say(driver, *args, "monitoring on")
    r = subprocess.run([os.path.join(BIN, driver), *args, f"--out={out}"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    mu, n, meta = _plot_load_tsv(out.read_text())
    assert meta == {"beta": 20.0, "t": 1.0, "tp": -0.35, "U": 5.75, "V": 0.9}, meta
    assert len(mu) == 221 and abs(mu[0] + 5.0) < 1e-12 and abs(mu[-1] - 25.0) < 1e-9
    assert np.all(np.diff(mu) > 0) and np.all((n >= 0) & (n <= 2.0))
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("METRIC")]
    assert len(lines) == 1
    three = _METRIC_THREE.match(lines[0].strip())
    assert three and three.group("kind") == ("FTLM_nmu_rect_k" if driver.endswith("_k") else "FTLM_nmu_rect")
    assert float(three.group("wall")) > 0 and float(three.group("rss_b")) > 0
    assert abs(float(three.group("rss_mib")) - float(three.group("rss_b")) / 2 ** 20) < 1.0
    if driver == "ftlm_nmu_rect_k":
        four = _METRIC_FOUR.match(lines[0].strip())
        assert four and four.group("solver") == "k-orbit"
        assert 0 < float(four.group("wall_sec")) <= float(four.group("wall"))


@pytest.mark.parametrize("lx,ly,nu,nd,kx,ky,dk", [
    (2, 2, 1, 1, 0, 0, 4), (2, 2, 1, 1, 1, 0, 4), (2, 2, 1, 1, 1, 1, 4), (2, 2, 2, 2, 0, 0, 12),
    (2, 2, 2, 2, 1, 0, 8), (2, 2, 0, 2, 0, 0, 0), (3, 2, 1, 1, 0, 0, 6), (3, 2, 1, 1, 1, 0, 6)])
def test_k_dim_analysis_cases(oracle, lx, ly, nu, nd, kx, ky, dk):
    """tests/test_k_dim_analysis.cpp:372-398 (t=1, t'=-0.2, U=4, V=0.5): on the reference's cases the
    device Gram block has dimension rank(P_K) of the dense translation projector, and its spectrum
    equals that of H restricted to range(P_K) (the reference's bound 1e-8; here 1e-10)."""
    import symmetry_kat as kat

    say("k-dim analysis", lx, ly, nu, nd, kx, ky)
    p = api.HubbardParams(lx, ly, 1.0, -0.2, 4.0, 0.5, 0.0, 0.0)
    g = api.KGramOperator.from_params(p, nu, nd, api.MomentumSector(kx, ky, lx, ly))
    Q = kat.projector_range(oracle, lx, ly, nu, nd, kx, ky)
    assert g.dim() == Q.shape[1] == dk
    if dk:
        H = oracle.dense_hamiltonian(params8(lx, ly, 1.0, -0.2, 4.0, 0.5), nu, nd)
        want = np.linalg.eigvalsh(Q.conj().T @ H @ Q)
        got = np.linalg.eigvalsh(_dense_from(g))
        assert np.max(np.abs(got - want)) < 1e-10


def test_momentum_logz_2x2(oracle):
    """tests/test_momentum_logz_2x2.cpp:96-166 (2x2, t=1, t'=-0.35, U=5.75, V=0.9, beta=20): the
    device ED of every full sector (the reference's real embedding and capped Jacobi) is finite and
    equals ln tr e^{-beta H} from numpy's eigh of the oracle's H (1e-10 relative); the Gram blocks
    are not all empty, and where they keep every state their ED log-sum equals the sector's."""
    say("momentum logZ 2x2")
    p = api.HubbardParams(2, 2, 1.0, -0.35, 5.75, 0.9, 0.0, 0.0)
    p8 = params8(2, 2, 1.0, -0.35, 5.75, 0.9)
    beta = np.array([20.0])
    nonempty = full_cover = 0
    for nu in range(5):
        for nd in range(5):
            sec = api.HubbardOperator.from_params(p, nu, nd)
            lz = float(api.exact_log_partition(sec, beta)[0])
            E = np.linalg.eigvalsh(oracle.dense_hamiltonian(p8, nu, nd))
            want = -20.0 * E[0] + math.log(np.sum(np.exp(-20.0 * (E - E[0]))))
            assert math.isfinite(lz) and abs(lz - want) <= 1e-10 * max(1.0, abs(want)), (nu, nd, lz, want)
            blocks = [api.KGramOperator.from_params(p, nu, nd, api.MomentumSector(kx, ky, 2, 2), sector=sec)
                      for ky in range(2) for kx in range(2)]
            dks = [b.dim() for b in blocks]
            nonempty += sum(dks) > 0
            if sum(dks) == sec.dim():
                full_cover += 1
                tot = -math.inf
                for b in blocks:
                    if b.dim():
                        tot = float(np.logaddexp(tot, api.exact_log_partition(b, beta)[0]))
                assert abs(tot - lz) <= 1e-9 * max(1.0, abs(lz)), (nu, nd, tot, lz)
    assert nonempty == 21 and full_cover == 21
