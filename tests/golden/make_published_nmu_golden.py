"""Writes tests/golden/nmu_published.json from the reference's published n(mu) runs of its
non-momentum driver (bench_ftlm_nmu_rect, defaults R=12, M=72, seed 7, beta=20, t=1, t'=-0.35,
U=5.75, V=0.9) on three lattices.

Sources (benchmarks/data): n_vs_mu_{4x2,3x3,2x5}_beta20_ftlm.tsv (the n(mu) table, with the run's
wall time in its last comment line) and n_vs_mu_{4x2,3x3,2x5}_beta20_ftlm.stderr.log (per-sector
dim and logZ to 6 significant digits; the 2x5 log stops after 38 of its 121 sectors).  The mu grids
follow from the tables: -5..25 for 4x2 (the driver's default), -8..14 for 3x3 and 2x5, 221 points.
Run from the repository root with the reference checkout at /root/reference.
"""
from __future__ import annotations

import json
import os
import re

REF = "/root/reference/benchmarks/data"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {}
    for lat in ("4x2", "3x3", "2x5"):
        lx, ly = map(int, lat.split("x"))
        rows, wall = [], None
        for line in open(os.path.join(REF, f"n_vs_mu_{lat}_beta20_ftlm.tsv")):
            if line.startswith("#"):
                m = re.search(r"wall_time_s=(\S+)", line)
                if m:
                    wall = float(m.group(1))
            elif line.strip():
                rows.append([float(v) for v in line.split()])
        sectors = {}
        for line in open(os.path.join(REF, f"n_vs_mu_{lat}_beta20_ftlm.stderr.log")):
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
