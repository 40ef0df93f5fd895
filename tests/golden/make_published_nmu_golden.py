"""Writes tests/golden/nmu_published.json from the reference's published n(mu) runs of its
non-momentum driver (bench_ftlm_nmu_rect, defaults R=12, M=72, seed 7, beta=20, t=1, t'=-0.35,
U=5.75, V=0.9) on three lattices.

Sources (benchmarks/data): n_vs_mu_{4x2,3x3,2x5,2x2}_beta20_ftlm.tsv and
n_vs_mu_3x2_beta20_ftlm_nomom.tsv (the n(mu) tables, with the run's wall time in their last comment
line), and the per-sector dim and logZ (6 significant digits) of n_vs_mu_*_beta20_ftlm.stderr.log
and of section 1 of n_vs_mu_3x2_beta20_three_way_monitoring.txt (the 2x5 log stops after 38 of its
121 sectors).  The mu grids follow from the tables.
Run from the repository root with the reference checkout at /root/reference.
"""
from __future__ import annotations

import json
import os
import re

REF = "/root/reference/benchmarks/data"
HERE = os.path.dirname(os.path.abspath(__file__))


# lattice -> (n(mu) table, per-sector log, section of the log to read or None for all of it)
RUNS = {
    "4x2": ("n_vs_mu_4x2_beta20_ftlm.tsv", "n_vs_mu_4x2_beta20_ftlm.stderr.log", None),
    "3x3": ("n_vs_mu_3x3_beta20_ftlm.tsv", "n_vs_mu_3x3_beta20_ftlm.stderr.log", None),
    "2x5": ("n_vs_mu_2x5_beta20_ftlm.tsv", "n_vs_mu_2x5_beta20_ftlm.stderr.log", None),
    "2x2": ("n_vs_mu_2x2_beta20_ftlm.tsv", "n_vs_mu_2x2_beta20_ftlm.stderr.log", None),
    "3x2": ("n_vs_mu_3x2_beta20_ftlm_nomom.tsv", "n_vs_mu_3x2_beta20_three_way_monitoring.txt",
            ("--- 1) FTLM non-momentum", "--- 2)")),
}


def main():
    out = {}
    for lat, (tsv, log, section) in RUNS.items():
        lx, ly = map(int, lat.split("x"))
        rows, wall = [], None
        for line in open(os.path.join(REF, tsv)):
            if line.startswith("#"):
                m = re.search(r"wall_time_s=(\S+)", line)
                if m:
                    wall = float(m.group(1))
            elif line.strip():
                rows.append([float(v) for v in line.split()])
        sectors = {}
        text = open(os.path.join(REF, log)).read()
        if section:
            text = text.split(section[0])[1].split(section[1])[0]
        for line in text.splitlines():
            m = re.match(r"sector \((\d+),(\d+)\) dim=(\d+) logZ=(\S+)", line.strip())
            if m:
                sectors[f"{m.group(1)},{m.group(2)}"] = [int(m.group(3)), float(m.group(4))]
        mu_min, mu_max = rows[0][0], rows[-1][0]
        args = [f"--Lx={lx}", f"--Ly={ly}", f"--mu-min={mu_min:g}", f"--mu-max={mu_max:g}", f"--n-mu={len(rows)}"]
        out[lat] = {"args": args, "reference_wall_time_s": wall, "sectors": sectors, "n_mu": rows}
    path = os.path.join(HERE, "nmu_published.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, {k: (len(v["sectors"]), len(v["n_mu"]), v["reference_wall_time_s"]) for k, v in out.items()})


if __name__ == "__main__":
    main()
