"""The n(mu) drivers on the device against their fixtures (tests/golden/nmu.json):
ftlm_nmu_rect against the reference's own bench_ftlm_nmu_rect output, ftlm_nmu_rect_k
(--solver=k-orbit) against the reference's orbit blocks, FTLM rule and seed schedule.  The drivers
print 6 significant digits (half an ulp is up to 5e-6 relative), so values agree to 1e-5 relative."""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ftlm_extended_hubbard_amd", "bin")
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "nmu.json")))
SECTOR_RE = re.compile(r"^sector \((\d+),(\d+)\) .*logZ=(\S+)$")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    api = pytest.importorskip("ftlm_extended_hubbard_amd.api")
    if api.device_count() < 1:
        pytest.skip("no HIP device")


def run(name, args, tmp_path):
    out = tmp_path / "nmu.tsv"
    r = subprocess.run([os.path.join(BIN, name), *args, "--no-monitor=1", f"--out={out}"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    sectors = {}
    for line in r.stderr.splitlines():
        m = SECTOR_RE.match(line.strip())
        if m:
            sectors[f"{m.group(1)},{m.group(2)}"] = float(m.group(3))
    mu, n = [], []
    for line in out.read_text().splitlines():
        if line.startswith("#") or not line.strip():
            continue
        a, b = line.split("\t")
        mu.append(float(a))
        n.append(float(b))
    return np.array(mu), np.array(n), sectors, r.stdout


def close(a, b, rel=1e-5, absol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= rel * np.abs(b) + absol)


@pytest.mark.parametrize("case", range(len(FIX["rect"])))
def test_nmu_rect_matches_reference_driver(case, tmp_path):
    c = FIX["rect"][case]
    say("ftlm_nmu_rect", *c["args"])
    mu, n, sectors, stdout = run("ftlm_nmu_rect", c["args"], tmp_path)
    assert "METRIC kind=FTLM_nmu_rect" in stdout
    assert close(mu, c["mu"]) and close(n, c["n"]), (n, c["n"])
    assert sectors.keys() == c["sector_logZ"].keys()
    for k, v in c["sector_logZ"].items():
        assert close(sectors[k], v), (k, sectors[k], v)


def test_nmu_rect_k_orbit_matches_reference_blocks(tmp_path):
    c = FIX["rect_k"][0]
    say("ftlm_nmu_rect_k", *c["args"])
    mu, n, sectors, stdout = run("ftlm_nmu_rect_k", c["args"], tmp_path)
    assert "METRIC kind=FTLM_nmu_rect_k solver=k-orbit" in stdout
    assert close(mu, c["mu"]) and close(n, c["n"]), (n, c["n"])
    # sectors whose every state drops out of the Gram blocks print logZ=-inf, as the reference's do
    assert {k for k, v in sectors.items() if np.isfinite(v)} == c["sector_logZ"].keys()
    for k, v in c["sector_logZ"].items():
        assert close(sectors[k], v), (k, sectors[k], v)


def test_nmu_rect_k_orbit_stacked_equals_per_block(tmp_path):
    """--solver=k-orbit --orbit-all=1 with random-phase starts on 4x2: the orbit blocks of every
    sector through momentum stacks (--path=stack, one ftlm_hip_ftlm_ritz_seeded call per sector
    with the per-K seeds) against the same blocks one by one (--path=per-op).  The printed 6 digits
    of every sector's logZ and of n(mu) agree (the stacked dot sums differ in the last bits only)."""
    args = ["--Lx=4", "--Ly=2", "--beta=4", "--n-mu=21", "--ftlm-random=6", "--lanczos-steps=24",
            "--solver=k-orbit", "--orbit-all=1", "--start=phase", "--quad=converged", "--ed-cutoff=24"]
    say("ftlm_nmu_rect_k stacked vs per-op", *args)
    mu0, n0, sec0, _ = run("ftlm_nmu_rect_k", [*args, "--path=per-op"], tmp_path)
    mu1, n1, sec1, _ = run("ftlm_nmu_rect_k", [*args, "--path=stack"], tmp_path)
    assert sec0.keys() == sec1.keys() and len(sec0) == 81
    assert np.array_equal(mu0, mu1) and close(n0, n1, rel=2e-6)
    for k in sec0:
        assert close(sec1[k], sec0[k], rel=2e-6), (k, sec0[k], sec1[k])


def test_nmu_rect_k_ed_matches_gram_spectra(oracle, tmp_path):
    """--solver=ed (every Gram block by device dense ED, as the reference's ed solver) on 3x2: each
    sector's logZ is ln sum_K tr e^{-beta H_K} over the oracle's Gram blocks (numpy eigvalsh), which
    is ln tr e^{-beta H} of the whole sector wherever no orbit drops out."""
    from pyoracle import params8

    args = ["--Lx=3", "--Ly=2", "--beta=1.5", "--n-mu=11", "--t=1", "--tp=-0.3", "--U=5", "--V=0.7"]
    say("ftlm_nmu_rect_k ed", *args)
    _, _, sectors, stdout = run("ftlm_nmu_rect_k", [*args, "--solver=ed"], tmp_path)
    assert "solver=ed" in stdout
    p8 = params8(3, 2, 1.0, -0.3, 5.0, 0.7)
    assert len(sectors) == 49
    full = 0
    for k, v in sectors.items():
        nu, nd = map(int, k.split(","))
        w = np.concatenate([np.linalg.eigvalsh(oracle.kgram_block(p8, nu, nd, kx, ky))
                            for ky in range(2) for kx in range(3)])
        if len(w) == 0:
            assert v == -np.inf, k
            continue
        mx = np.max(-1.5 * w)
        exact = mx + np.log(np.sum(np.exp(-1.5 * w - mx)))
        assert close(v, exact), (k, v, exact)
        if len(w) == len(np.linalg.eigvalsh(oracle.dense_hamiltonian(p8, nu, nd))):
            full += 1
    assert full == 41  # 49 sectors less (0,2), (2,0), (2,4), (4,2), (4,6), (6,4), (0,6), (6,0)


NUM = r"(-?[0-9.]+(?:e[-+]?[0-9]+)?)"


def run_plain(name, args):
    r = subprocess.run([os.path.join(BIN, name), *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_bench_2x2_matches_reference_driver():
    """ftlm_bench_2x2 (bench_2x2.cpp): the 1600-rotation dense Jacobi extrema (same input matrix,
    so the same capped result to the printed digits), the dense-vs-apply check, the steps used.
    The 36-step Lanczos extrema (M = dim, no reorthogonalisation, 120-rotation Jacobi far from
    converged) depend on the ghost-dominated late alpha/beta, which differ between summation orders;
    each stage of that line is pinned on identical inputs by test_extrema_config1, so here the two
    printed values only agree to 1e-3 of the spectral width."""
    c = FIX["bench"]["bench_2x2"]
    say("ftlm_bench_2x2")
    txt = run_plain("ftlm_bench_2x2", ["--matvec-iters=2000"])
    m = re.search(r"E_min=" + NUM + r"  E_max=" + NUM, txt)
    assert close(float(m.group(1)), c["dense_min"]) and close(float(m.group(2)), c["dense_max"]), txt
    err = float(re.search(r"max\|H_dense v - apply_H v\| = " + NUM, txt).group(1))
    assert err < 1e-12, txt
    m = re.search(r"used=(\d+) .*E_min≈" + NUM + r"  E_max≈" + NUM, txt)
    assert int(m.group(1)) == c["lz_used"]
    width = c["dense_max"] - c["dense_min"]
    assert abs(float(m.group(2)) - c["lz_min"]) < 1e-3 * width, txt
    assert abs(float(m.group(3)) - c["lz_max"]) < 1e-3 * width, txt
    assert "Mapply/s" in txt


def test_bench_4x2_ftlm_matches_reference_driver():
    """ftlm_bench_4x2_ftlm (bench_4x2_ftlm.cpp): averaged Ritz extrema over samples seeded
    seed + s * 7919, steps used and H*v count."""
    c = FIX["bench"]["bench_4x2_ftlm"]
    say("ftlm_bench_4x2_ftlm", *c["args"])
    txt = run_plain("ftlm_bench_4x2_ftlm", c["args"])
    m = re.search(r"avg Ritz E_min=" + NUM + r"  E_max=" + NUM, txt)
    assert close(float(m.group(1)), c["avg_min"]) and close(float(m.group(2)), c["avg_max"]), txt
    assert int(re.search(r"sum_lanczos_steps_used=(\d+)", txt).group(1)) == c["sum_steps"]
    assert int(re.search(r"matvec_count=(\d+)", txt).group(1)) == c["matvec_count"]


def test_bench_2x2_nmu_matches_reference_driver():
    """ftlm_bench_2x2_nmu (bench_2x2_nmu.cpp): all 256 levels by device dense Jacobi, n(mu)."""
    c = FIX["bench"]["bench_2x2_nmu"]
    say("ftlm_bench_2x2_nmu", *c["args"])
    txt = run_plain("ftlm_bench_2x2_nmu", c["args"])
    assert int(re.search(r"exact diagonalization: (\d+) eigenlevels", txt).group(1)) == c["levels"]
    rows = [line.split("\t") for line in txt.splitlines() if "\t" in line and not line.startswith("#")]
    assert close([float(a) for a, _ in rows], c["mu"]) and close([float(b) for _, b in rows], c["n"]), rows


PUB = json.load(open(os.path.join(ROOT, "tests", "golden", "nmu_published.json")))


@pytest.mark.parametrize("lat", ["2x2", "3x2", "4x2", "3x3", "2x5"])
def test_nmu_rect_reproduces_published_run(lat, tmp_path):
    """The reference's published runs of bench_ftlm_nmu_rect with its defaults (benchmarks/data/
    n_vs_mu_*_beta20_ftlm* on 2x2, 3x2, 4x2, 3x3, 2x5): every logged sector's logZ and the n(mu) table, to the
    printed 6 digits, with the same seeds and capped-Jacobi rule.  Reports the device wall time next
    to the reference's published single-core time."""
    c = PUB[lat]
    say("ftlm_nmu_rect published", lat, *c["args"])
    import time

    t0 = time.perf_counter()
    mu, n, sectors, stdout = run("ftlm_nmu_rect", c["args"], tmp_path)
    wall = time.perf_counter() - t0
    print(f"[case] {lat}: device run {wall:.2f} s (reference {c['reference_wall_time_s']} s)", flush=True)
    ref = np.array(c["n_mu"])
    assert close(mu, ref[:, 0], rel=1e-5, absol=1e-4)
    assert close(n, ref[:, 1], rel=2e-5, absol=1e-12), np.max(np.abs(n - ref[:, 1]))
    bad = [(k, sectors[k], v[1]) for k, v in c["sectors"].items() if not close(sectors[k], v[1], rel=6e-6)]
    assert not bad, bad
