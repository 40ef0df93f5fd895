"""Write the golden fixtures in tests/golden/ from the reference itself.

Run in a container where /root/reference exists:  python tests/golden/make_golden.py
It uses oracle/_ref/libftlm_ref.so (the reference's hot-path sources compiled by oracle/Makefile)
for computed vectors, and parses two published data files of the reference
(benchmarks/data/eigs_full_vs_k_2x2.tsv, benchmarks/data/n_vs_mu_beta20.tsv) into arrays.
The fixtures are data only (inputs + expected outputs); nothing of the reference's source travels.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from pyoracle import Reference, build, params8  # noqa: E402

REF_DATA = "/root/reference/benchmarks/data"


def seeded_vec(dim: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    return x / np.linalg.norm(x)


# (name, Lx, Ly, t, tp, U, V, phi_x, phi_y)
MODELS = {
    "c1_2x2": (2, 2, 1.0, -0.2, 4.0, 0.5, 0.0, 0.0),
    "pub_2x2": (2, 2, 1.0, -0.35, 5.75, 0.9, 0.0, 0.0),
    "flux_2x2": (2, 2, 1.0, 0.0, 0.0, 0.0, 0.31, -0.27),
    "c2_4x2": (4, 2, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0),
    "flux_3x2": (3, 2, 1.0, -0.3, 5.0, 0.7, 0.2, 0.1),
    "c5_4x3": (4, 3, 1.0, -0.25, 8.0, 1.0, 0.05, 0.0),
    "c3_4x4": (4, 4, 1.0, -0.25, 8.0, 1.5, 0.1, 0.0),
}

# (model, nu, nd) sectors for H*v / Lanczos fixtures
APPLY_SECTORS = [
    ("c1_2x2", 2, 2), ("pub_2x2", 1, 2), ("flux_2x2", 2, 2), ("c2_4x2", 4, 4), ("c2_4x2", 3, 1),
    ("flux_3x2", 3, 2), ("c5_4x3", 2, 2), ("c3_4x4", 2, 1),
]


def main() -> None:
    build(with_ref=True)
    ref = Reference()
    out = {}
    for name, m in MODELS.items():
        hops, pairs = ref.geometry(params8(*m))
        out[f"geom/{name}/model"] = np.array(m, np.float64)
        out[f"geom/{name}/hops_from"] = hops["from"]
        out[f"geom/{name}/hops_to"] = hops["to"]
        out[f"geom/{name}/hops_spin"] = hops["spin"]
        out[f"geom/{name}/hops_coeff"] = hops["re"] + 1j * hops["im"]
        out[f"geom/{name}/pairs"] = pairs

    # basis masks + index_of (in-sector and out-of-sector probes)
    for (n, nu, nd) in [(4, 2, 2), (4, 0, 3), (8, 4, 4), (8, 3, 1), (12, 2, 1), (16, 1, 2), (6, 3, 3)]:
        up, dn = ref.basis_masks(n, nu, nd)
        key = f"basis/{n}_{nu}_{nd}"
        out[key + "/up"] = up
        out[key + "/dn"] = dn
        rng = np.random.default_rng(n * 100 + nu * 10 + nd)
        pu = rng.integers(0, 1 << n, size=64, dtype=np.uint64)
        pd = rng.integers(0, 1 << n, size=64, dtype=np.uint64)
        pick = rng.integers(0, len(up), size=64)
        pu = np.concatenate([pu, up[pick]])
        pd = np.concatenate([pd, dn[pick]])
        out[key + "/probe_up"] = pu
        out[key + "/probe_dn"] = pd
        out[key + "/probe_idx"] = ref.index_of(n, nu, nd, pu, pd)

    for (name, nu, nd) in APPLY_SECTORS:
        m = MODELS[name]
        p8 = params8(*m)
        n = m[0] * m[1]
        up, dn = ref.basis_masks(n, nu, nd)
        dim = len(up)
        x = seeded_vec(dim, 1000 + 10 * nu + nd)
        key = f"apply/{name}/{nu}_{nd}"
        out[key + "/x"] = x
        out[key + "/y"] = ref.apply(p8, nu, nd, x)
        for seed in (7, 42):
            M = min(40, 2 * dim + 2)
            used, a, b = ref.lanczos(p8, nu, nd, M, seed)
            out[f"lanczos/{name}/{nu}_{nd}/s{seed}/M"] = np.array([M])
            out[f"lanczos/{name}/{nu}_{nd}/s{seed}/alpha"] = a
            out[f"lanczos/{name}/{nu}_{nd}/s{seed}/beta"] = b
            out[f"lanczos/{name}/{nu}_{nd}/s{seed}/used"] = np.array([used])

    # config-1 harness values: lanczos_extrema (capped 120-rotation Jacobi), 36 steps, seed 7
    used, lo, hi = ref.lanczos_extrema(params8(*MODELS["c1_2x2"]), 2, 2, 36, 7)
    out["extrema/c1_2x2/2_2/s7"] = np.array([used, lo, hi])
    used, lo, hi = ref.lanczos_extrema(params8(2, 1, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0), 1, 1, 48, 42)
    out["extrema/chain_2x1/1_1/s42"] = np.array([used, lo, hi])

    # per-sector FTLM lnZ (complex and real variants), 4x2 all 81 sectors, small R/M
    p8 = params8(*MODELS["c2_4x2"])
    lnz_c, lnz_r = np.zeros((9, 9)), np.zeros((9, 9))
    for nu in range(9):
        for nd in range(9):
            lnz_c[nu, nd] = ref.ftlm_log_partition(p8, nu, nd, 1.0, 2, 16, 7, False)
            lnz_r[nu, nd] = ref.ftlm_log_partition(p8, nu, nd, 1.0, 2, 16, 7, True)
    out["ftlm/c2_4x2/beta1_R2_M16_s7/complex"] = lnz_c
    out["ftlm/c2_4x2/beta1_R2_M16_s7/real"] = lnz_r
    # long Lanczos where the capped Jacobi bites (4x2 (4,4), M=80/150, one sample)
    for M in (80, 150):
        for bt in (0.5, 2.0, 8.0):
            out[f"ftlm/c2_4x2/4_4/M{M}_beta{bt}"] = np.array(
                [ref.ftlm_log_partition(p8, 4, 4, bt, 1, M, 7, False)])
    # the reference's own thermo test operator diag(2,3), beta=1, R=48, M=32, seed 12345
    out["ftlm/diag23"] = np.array([ref.ftlm_dense(np.diag([2.0, 3.0]).astype(np.complex128), 1.0, 48, 32, 12345)])

    # orbit k-blocks
    for (lx, ly, nu, nd) in [(2, 2, 2, 2), (3, 2, 3, 3), (4, 2, 4, 4), (4, 2, 2, 3)]:
        dims = np.zeros((ly, lx), np.int64)
        for ky in range(ly):
            for kx in range(lx):
                reps, osz, ssz, A = ref.orbit_basis(lx, ly, nu, nd, kx, ky)
                dims[ky, kx] = len(reps)
                if (lx, ly) != (4, 2) or (kx, ky) in ((1, 0), (0, 1)):
                    key = f"orbit/{lx}x{ly}/{nu}_{nd}/k{kx}{ky}"
                    out[key + "/reps"] = reps
                    out[key + "/orbit_size"] = osz
                    out[key + "/stab_size"] = ssz
                    out[key + "/A"] = A
        out[f"orbit/{lx}x{ly}/{nu}_{nd}/dims"] = dims
    for (lx, ly, nu, nd, kx, ky) in [(3, 2, 3, 3, 1, 0), (2, 2, 2, 2, 0, 0), (4, 2, 4, 4, 2, 1), (4, 2, 2, 3, 1, 0)]:
        p8 = params8(lx, ly, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0)
        reps = ref.orbit_basis(lx, ly, nu, nd, kx, ky)[0]
        x = seeded_vec(len(reps), 77 + kx + 3 * ky)
        key = f"orbit_apply/{lx}x{ly}/{nu}_{nd}/k{kx}{ky}"
        out[key + "/x"] = x
        out[key + "/y"] = ref.orbit_apply(p8, nu, nd, kx, ky, x)
        up, dn = ref.basis_masks(lx * ly, nu, nd)
        st = ((up.astype(np.uint32) << np.uint32(16)) | dn.astype(np.uint32)).astype(np.uint32)
        row, dx, dy, eta = ref.orbit_map(lx, ly, nu, nd, kx, ky, st)
        out[key + "/map"] = np.stack([row, dx, dy, eta.astype(np.int32)], axis=1)
        out[key + "/lnZ_b1.5_R2_M12_s3"] = np.array([ref.orbit_ftlm_log_partition(p8, nu, nd, kx, ky, 1.5, 2, 12, 3)])

    # published reference data files
    parse_eigs(out)
    parse_n_vs_mu(out)

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), len(out), "arrays")


def parse_eigs(out: dict) -> None:
    """benchmarks/data/eigs_full_vs_k_2x2.tsv: sorted full-sector eigenvalues for all 2x2 sectors
    at t=1, t'=-0.35, U=5.75, V=0.9 (rows of the eigs_pairwise table)."""
    rows = []
    section = None
    with open(os.path.join(REF_DATA, "eigs_full_vs_k_2x2.tsv")) as f:
        for line in f:
            line = line.strip()
            if line.startswith("# eigs_pairwise"):
                section = "pair"
                continue
            if line.startswith("#"):
                section = None if not line.startswith("# eigs_pairwise") else section
                continue
            if section == "pair" and line and not line.startswith("n_up"):
                parts = line.split("\t")
                rows.append([float(v) for v in parts[:5]])
    arr = np.array(rows)
    out["published/eigs_2x2/rows"] = arr  # n_up, n_dn, dim, idx, eval_full_sorted


def parse_n_vs_mu(out: dict) -> None:
    """benchmarks/data/n_vs_mu_beta20.tsv: 2x2 PBC grand-canonical ED n(mu), beta=20."""
    vals = []
    with open(os.path.join(REF_DATA, "n_vs_mu_beta20.tsv")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            a, b = line.split()
            vals.append([float(a), float(b)])
    out["published/n_vs_mu_2x2_beta20"] = np.array(vals)


if __name__ == "__main__":
    main()
