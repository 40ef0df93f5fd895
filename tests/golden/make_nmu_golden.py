"""Write tests/golden/nmu.json: expected output of the grand-canonical n(mu) drivers.

Run where /root/reference exists (after `make -C oracle ref refbench`):
    python tests/golden/make_nmu_golden.py
- "rect": the reference's own bench_ftlm_nmu_rect (benchmarks/bench_ftlm_nmu_rect.cpp, built from
  its sources into oracle/_ref/) on small lattices: the n(mu) table and every sector's logZ.
- "bench": the printed numbers of the reference's bench_2x2, bench_4x2_ftlm and bench_2x2_nmu
  (built the same way).
- "rect_k": the momentum-block solver of benchmarks/bench_ftlm_nmu_4x4_k.cpp with
  --solver=k-orbit on a 3x2 lattice.  That driver needs LAPACK and is not built here; its
  per-block rule (:1072-1079: ED below the cutoff, else ftlm_log_partition_complex on
  apply_orbit_k_block with seed + idx*257 + ky*17 + kx) is evaluated with the reference's orbit
  basis / apply / FTLM (oracle/_ref) and the oracle's dense ED, then combined as the driver does.
  Sectors outside its orbit-direct gate (N_up or N_dn in {0, N}, or more than 512 states,
  :800-807) use the Gram block (:1003-1019), restated by the oracle (kgram_block); on 3x2 all
  of them are below the ED cutoff, so their logZ is the oracle's ED of that block.
Fixtures are data only (arguments, n(mu), logZ per sector).
"""
from __future__ import annotations

import json
import math
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from pyoracle import Oracle, Reference, params8  # noqa: E402

REF_NMU = os.path.join(ROOT, "oracle", "_ref", "bench_ftlm_nmu_rect")

RECT_CASES = [
    ["--Lx=2", "--Ly=2", "--beta=3", "--n-mu=25", "--ftlm-random=4", "--lanczos-steps=20"],
    ["--Lx=3", "--Ly=2", "--beta=2", "--n-mu=31", "--ftlm-random=3", "--lanczos-steps=24", "--t=1", "--tp=-0.3",
     "--U=5", "--V=0.7"],
    ["--Lx=4", "--Ly=2", "--beta=4", "--n-mu=41", "--ftlm-random=2", "--lanczos-steps=30", "--t=1", "--tp=-0.25",
     "--U=8", "--V=1"],
]
K_CASE = dict(Lx=3, Ly=2, beta=2.0, mu_min=-5.0, mu_max=25.0, n_mu=31, n_rand=3, lz=24, seed=7, ed_cutoff=64,
              t=1.0, tp=-0.35, U=5.75, V=0.9)

SECTOR_RE = re.compile(r"^sector \((\d+),(\d+)\) dim=(\d+) logZ=(\S+)$")


def run_rect(args):
    out = "/tmp/_nmu_golden.tsv"
    r = subprocess.run([REF_NMU, *args, "--no-monitor=1", f"--out={out}"], capture_output=True, text=True, check=True)
    sectors = {}
    for line in r.stderr.splitlines():
        m = SECTOR_RE.match(line.strip())
        if m:
            sectors[f"{m.group(1)},{m.group(2)}"] = float(m.group(4))
    mu, n = [], []
    for line in open(out):
        if line.startswith("#") or not line.strip():
            continue
        a, b = line.split("\t")
        mu.append(float(a))
        n.append(float(b))
    return {"args": args, "mu": mu, "n": n, "sector_logZ": sectors}


def nmu_table(logZ: dict, n_sites: int, beta: float, mu_min: float, mu_max: float, n_mu: int):
    """bench_ftlm_nmu_rect.cpp:186-219."""
    mus, ns = [], []
    for k in range(n_mu):
        mu = mu_min + (k / (n_mu - 1)) * (mu_max - mu_min)
        ex = {key: lz + beta * mu * sum(map(int, key.split(","))) for key, lz in logZ.items() if math.isfinite(lz)}
        mx = max(ex.values())
        sw = sum(math.exp(v - mx) for v in ex.values())
        snw = sum(sum(map(int, key.split(","))) * math.exp(v - mx) for key, v in ex.items())
        mus.append(mu)
        ns.append(snw / sw / n_sites)
    return mus, ns


def run_rect_k(c):
    ref, orc = Reference(), Oracle()
    lx, ly = c["Lx"], c["Ly"]
    n = lx * ly
    p8 = params8(lx, ly, c["t"], c["tp"], c["U"], c["V"], 0.0, 0.0)
    logZ, blocks = {}, {}
    for nu in range(n + 1):
        for nd in range(n + 1):
            idx = nu * (n + 1) + nd
            dim_full = math.comb(n, nu) * math.comb(n, nd)
            # orbit coordinates only for interior sectors of <= 512 states; HubbardMomentumBlock
            # (Gram basis) for the others (bench_ftlm_nmu_4x4_k.cpp:800-807, 1003-1019)
            orbit = 0 < nu < n and 0 < nd < n and dim_full <= 512
            lzs = []
            for ky in range(ly):
                for kx in range(lx):
                    if orbit:
                        dk = len(ref.orbit_basis(lx, ly, nu, nd, kx, ky)[0])
                    else:
                        Hg = orc.kgram_block(p8, nu, nd, kx, ky)
                        dk = Hg.shape[0]
                    if dk <= 0:
                        continue
                    if dk <= c["ed_cutoff"]:
                        if orbit:
                            H = np.zeros((dk, dk), np.complex128)
                            for j in range(dk):
                                e = np.zeros(dk, np.complex128)
                                e[j] = 1.0
                                H[:, j] = ref.orbit_apply(p8, nu, nd, kx, ky, e)
                        else:
                            H = Hg
                        lz = orc.exact_log_partition(H, c["beta"])
                    else:
                        # FTLM on a Gram block draws its starts in the reference's zheev basis,
                        # which no other build reproduces: keep the fixture to orbit FTLM blocks
                        assert orbit, (nu, nd, kx, ky)
                        seed_k = c["seed"] + idx * 257 + ky * 17 + kx
                        lz = ref.orbit_ftlm_log_partition(p8, nu, nd, kx, ky, c["beta"], c["n_rand"], c["lz"], seed_k)
                    lzs.append(lz)
                    blocks[f"{nu},{nd},{kx},{ky}"] = [dk, lz]
            if lzs:
                mx = max(lzs)
                logZ[f"{nu},{nd}"] = mx + math.log(sum(math.exp(v - mx) for v in lzs))
    mu, nn = nmu_table(logZ, n, c["beta"], c["mu_min"], c["mu_max"], c["n_mu"])
    args = [f"--Lx={lx}", f"--Ly={ly}", f"--beta={c['beta']:g}", f"--n-mu={c['n_mu']}", f"--ftlm-random={c['n_rand']}",
            f"--lanczos-steps={c['lz']}", f"--seed={c['seed']}", f"--ed-cutoff={c['ed_cutoff']}", "--solver=k-orbit"]
    return {"args": args, "mu": mu, "n": nn, "sector_logZ": logZ, "blocks": blocks}


NUM = r"(-?[0-9.]+(?:e[-+]?[0-9]+)?)"
BENCH_CASES = {
    "bench_2x2": ["--matvec-iters=0"],
    "bench_4x2_ftlm": ["--samples=8", "--lanczos-steps=64"],
    "bench_2x2_nmu": ["--beta=5", "--n-mu=31"],
}


def run_bench(name, args):
    """The printed numbers of the reference's small drivers (bench_2x2, bench_4x2_ftlm, bench_2x2_nmu)."""
    r = subprocess.run([os.path.join(ROOT, "oracle", "_ref", name), *args], capture_output=True, text=True,
                       check=True)
    txt = r.stdout
    out = {"args": args}
    if name == "bench_2x2":
        m = re.search(r"E_min=" + NUM + r"  E_max=" + NUM, txt)
        out["dense_min"], out["dense_max"] = float(m.group(1)), float(m.group(2))
        m = re.search(r"used=(\d+) .*E_min≈" + NUM + r"  E_max≈" + NUM, txt)
        out["lz_used"], out["lz_min"], out["lz_max"] = int(m.group(1)), float(m.group(2)), float(m.group(3))
    elif name == "bench_4x2_ftlm":
        m = re.search(r"avg Ritz E_min=" + NUM + r"  E_max=" + NUM, txt)
        out["avg_min"], out["avg_max"] = float(m.group(1)), float(m.group(2))
        out["sum_steps"] = int(re.search(r"sum_lanczos_steps_used=(\d+)", txt).group(1))
        out["matvec_count"] = int(re.search(r"matvec_count=(\d+)", txt).group(1))
    else:
        rows = [line.split("\t") for line in txt.splitlines() if "\t" in line and not line.startswith("#")]
        out["mu"] = [float(a) for a, _ in rows]
        out["n"] = [float(b) for _, b in rows]
        out["levels"] = int(re.search(r"exact diagonalization: (\d+) eigenlevels", txt).group(1))
    return out


def main():
    out = {"rect": [run_rect(a) for a in RECT_CASES], "rect_k": [run_rect_k(K_CASE)],
           "bench": {k: run_bench(k, v) for k, v in BENCH_CASES.items()}}
    path = os.path.join(HERE, "nmu.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
