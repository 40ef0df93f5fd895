"""GPU: the drop-in C++ API.  Runs the C++ test programs of tests/cpp (written against the
reference's headers) and checks the pybind11 module against the golden vectors."""

import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prog", ["test_pipeline", "test_thermo", "test_orbit", "test_apply_host"])
def test_cpp_program(prog):
    path = os.path.join(BIN, prog)
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run __graft_entry__.build()")
    say("cpp", prog)
    r = subprocess.run([path], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert f"{prog} ok" in r.stdout


@pytest.mark.parametrize("prog", ["test_momentum_partition_2x2", "test_orbit_k_block_basis"])
def test_reference_symmetry_tests_on_device(prog):
    """The reference's own tests (tests/*.cpp of the reference checkout), compiled against the
    drop-in headers and libftlm.so into oracle/_ref/tests where that checkout exists; these two
    build FockBasis / OrbitKBlockBasis, which live on the device."""
    path = os.path.join(ROOT, "oracle", "_ref", "tests", prog)
    if not os.path.exists(path):
        pytest.skip("reference tests not built here (no reference checkout when the tree was built)")
    say("reference test", prog)
    r = subprocess.run([path], capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    assert "ok" in r.stdout


def test_pybind_hot_path(golden):
    from ftlm_extended_hubbard_amd import _ftlm_core as c

    say("pybind apply/lanczos c1_2x2")
    m = golden["geom/c1_2x2/model"]
    p = c.HubbardParams(int(m[0]), int(m[1]), *map(float, m[2:]))
    hops, pairs = c.build_hubbard_geometry(p, c.RectLattice(p.Lx, p.Ly))
    b = c.FockBasis(p.Lx * p.Ly, 2, 2)
    assert b.dim() == 36
    x = golden["apply/c1_2x2/2_2/x"]
    y = c.apply_extended_hubbard(p, b, hops, pairs, x)
    want = golden["apply/c1_2x2/2_2/y"]
    assert np.max(np.abs(y - want)) < 1e-12 * np.max(np.abs(want))
    op = c.SectorOperator(p, b, hops, pairs)
    used, a, bt = op.lanczos_tridiagonal(36, 7)
    wa = golden["lanczos/c1_2x2/2_2/s7/alpha"]
    n = min(12, len(wa))
    assert np.max(np.abs(np.array(a[:n]) - wa[:n])) < 1e-10


def test_pybind_gram_blocks():
    """_ftlm_core.GramBlockOperator (HubbardMomentumBlock counterpart): the 3x2 (2,4) blocks keep
    216 of 225 states, as the reference's k-gram run prints, and apply Hermitian blocks."""
    from ftlm_extended_hubbard_amd import _ftlm_core as c

    say("pybind gram blocks 3x2 (2,4)")
    p = c.HubbardParams(3, 2, 1.0, -0.35, 5.75, 0.9)
    hops, pairs = c.build_hubbard_geometry(p, c.RectLattice(3, 2))
    fb = c.FockBasis(6, 2, 4)
    full = c.SectorOperator(p, fb, hops, pairs)
    total = 0
    for ky in range(2):
        for kx in range(3):
            g = c.GramBlockOperator(full, p, fb, c.MomentumSector(kx, ky, 3, 2), hops, pairs)
            d = g.dim()
            total += d
            H = np.array([g.apply(np.eye(d, dtype=complex)[:, j]) for j in range(d)]).T
            assert np.max(np.abs(H - H.conj().T)) < 1e-12 * max(1.0, np.max(np.abs(H)))
    assert total == 216
