"""Writes tests/golden/nmu3x2_solvers.json from the reference's own 3x2 four-solver run.

Source: benchmarks/data/n_vs_mu_3x2_beta20_four_solver_monitoring.txt, the console output of
benchmarks/run_3x2_four_solver_benchmark.sh (bench_ftlm_nmu_rect_k --Lx=3 --Ly=2 --beta=20
--mu-min=-5 --mu-max=25 --n-mu=221 --ftlm-random=12 --lanczos-steps=72 --seed=7 with
--solver=k-gram / k-orbit / ed), and the n(mu) tables benchmarks/data/
n_vs_mu_3x2_beta20_solver_{k_gram,k_orbit,ed}.tsv it wrote.  That binary needs LAPACK (zheev)
and is not built here; the fixture is its printed data: per solver and sector dim_full,
dim_k_total, ED / FTLM block counts and logZ (6 significant digits), and the n(mu) rows.  Run from
the repository root with the reference checkout at /root/reference.
"""
from __future__ import annotations

import json
import os
import re

REF = "/root/reference/benchmarks/data"
HERE = os.path.dirname(os.path.abspath(__file__))
SECTIONS = {"k-gram": ("--- 2) solver=k-gram", "n_vs_mu_3x2_beta20_solver_k_gram.tsv"),
            "k-orbit": ("--- 3) solver=k-orbit", "n_vs_mu_3x2_beta20_solver_k_orbit.tsv"),
            "ed": ("--- 4) solver=ed", "n_vs_mu_3x2_beta20_solver_ed.tsv")}
ARGS = ["--Lx=3", "--Ly=2", "--beta=20", "--mu-min=-5", "--mu-max=25", "--n-mu=221", "--ftlm-random=12",
        "--lanczos-steps=72", "--seed=7", "--no-log-k-dims=1"]


def main():
    text = open(os.path.join(REF, "n_vs_mu_3x2_beta20_four_solver_monitoring.txt")).read()
    out = {"source": "benchmarks/data/n_vs_mu_3x2_beta20_four_solver_monitoring.txt and the solver TSVs",
           "args": ARGS, "solvers": {}}
    for name, (marker, tsv) in SECTIONS.items():
        part = text.split(marker)[1].split("\n--- ")[0]
        sectors = {}
        for m in re.finditer(r"sector \((\d+),(\d+)\) dim_full=(\d+) dim_k_total=(\d+) ed_blocks=(\d+) "
                             r"ftlm_blocks=(\d+) logZ=(\S+)", part):
            nu, nd, df, dk, ne, nf, lz = m.groups()
            sectors[f"{nu},{nd}"] = [int(df), int(dk), int(ne), int(nf), float(lz)]
        metric = re.search(r"METRIC kind=FTLM_nmu_rect_k solver=(\S+)", part).group(1)
        rows = [[float(v) for v in line.split()] for line in open(os.path.join(REF, tsv))
                if line.strip() and not line.startswith("#")]
        assert len(sectors) == 49 and len(rows) == 221 and metric == name
        out["solvers"][name] = {"sectors": sectors, "n_mu": rows}
    path = os.path.join(HERE, "nmu3x2_solvers.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
