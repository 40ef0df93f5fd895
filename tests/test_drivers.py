"""The n(mu) drivers without a device: built, flag handling, and the k-orbit fixture pinned
against exact diagonalisation (the union of the momentum blocks' spectra is the sector's)."""
from __future__ import annotations

import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ftlm_extended_hubbard_amd", "bin")
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from pyoracle import params8  # noqa: E402


def test_driver_binaries_and_solver_flags():
    for name in ("ftlm_nmu_rect", "ftlm_nmu_rect_k"):
        assert os.access(os.path.join(BIN, name), os.X_OK), name
    r = subprocess.run([os.path.join(BIN, "ftlm_nmu_rect_k"), "--solver=nope"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown --solver" in r.stderr
    assert "k-gram" in r.stderr and "k-orbit" in r.stderr
    r = subprocess.run([os.path.join(BIN, "ftlm_nmu_rect_k"), "--Lx=5", "--Ly=4"], capture_output=True, text=True)
    assert r.returncode == 2 and "Lx*Ly <= 16" in r.stderr


def test_k_orbit_fixture_against_exact_spectra(oracle):
    """Every sector of the 3x2 k-orbit fixture whose blocks all went through ED and cover the whole
    sector: the blocks' log-sum-exp equals ln tr e^{-beta H} of the full sector (numpy eigvalsh).
    Sectors whose Gram blocks drop orbits with an odd fermionic stabiliser sign (as the
    reference's do) cover fewer states: there the dims must match the oracle's Gram blocks."""
    d = json.load(open(os.path.join(ROOT, "tests", "golden", "nmu.json")))["rect_k"][0]
    args = dict(a[2:].split("=") for a in d["args"])
    lx, ly, beta = int(args["Lx"]), int(args["Ly"]), float(args["beta"])
    p8 = params8(lx, ly, 1.0, -0.35, 5.75, 0.9)
    checked = 0
    n = lx * ly
    for key, lz in d["sector_logZ"].items():
        nu, nd = map(int, key.split(","))
        dims = [v[0] for k, v in d["blocks"].items() if k.startswith(f"{nu},{nd},")]
        if max(dims) > int(args["ed-cutoff"]):
            continue
        if sum(dims) != math.comb(n, nu) * math.comb(n, nd):
            gram = sum(oracle.kgram_block(p8, nu, nd, kx, ky).shape[0] for ky in range(ly) for kx in range(lx))
            assert sum(dims) == gram, (key, dims, gram)
            continue
        w = np.linalg.eigvalsh(oracle.dense_hamiltonian(p8, nu, nd))
        mx = np.max(-beta * w)
        exact = mx + math.log(np.sum(np.exp(-beta * w - mx)))
        assert abs(lz - exact) <= 1e-9 * max(1.0, abs(exact)), (key, lz, exact)
        checked += 1
    assert checked >= 40
