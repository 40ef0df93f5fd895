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
    assert sectors.keys() == c["sector_logZ"].keys()
    for k, v in c["sector_logZ"].items():
        assert close(sectors[k], v), (k, sectors[k], v)


def test_nmu_rect_k_ed_matches_exact_spectra(oracle, tmp_path):
    """--solver=ed (every block by device dense ED) on 3x2: each sector's logZ is ln tr e^{-beta H}
    of the whole sector (numpy eigvalsh of the oracle's dense H)."""
    from pyoracle import params8

    args = ["--Lx=3", "--Ly=2", "--beta=1.5", "--n-mu=11", "--t=1", "--tp=-0.3", "--U=5", "--V=0.7"]
    say("ftlm_nmu_rect_k ed", *args)
    _, _, sectors, stdout = run("ftlm_nmu_rect_k", [*args, "--solver=ed"], tmp_path)
    assert "solver=ed" in stdout
    p8 = params8(3, 2, 1.0, -0.3, 5.0, 0.7)
    assert len(sectors) == 49
    for k, v in sectors.items():
        nu, nd = map(int, k.split(","))
        w = np.linalg.eigvalsh(oracle.dense_hamiltonian(p8, nu, nd))
        mx = np.max(-1.5 * w)
        exact = mx + np.log(np.sum(np.exp(-1.5 * w - mx)))
        assert close(v, exact), (k, v, exact)
