"""The drop-in C++ API (libftlm.so) and its pybind11 module (_ftlm_core), on the CPU: the binding
surface the reference's python/bindings.cpp:11-38 defines and the C++ entry points the reference's
headers declare (no device calls)."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ftlm_extended_hubbard_amd")


def core():
    try:
        from ftlm_extended_hubbard_amd import _ftlm_core
    except ImportError as e:  # not built in this checkout
        pytest.skip(f"_ftlm_core not built: {e}")
    return _ftlm_core


def test_hubbard_params_binding():
    c = core()
    p = c.HubbardParams()
    assert (p.Lx, p.Ly, p.t, p.tp, p.U, p.V, p.phi_x, p.phi_y) == (2, 2, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    q = c.HubbardParams(Lx=4, Ly=3, t=1.0, tp=-0.25, U=8.0, V=1.0, phi_x=0.05, phi_y=0.0)
    assert (q.Lx, q.Ly, q.tp, q.U, q.V, q.phi_x) == (4, 3, -0.25, 8.0, 1.0, 0.05)
    q.U = 4.0
    assert q.U == 4.0
    assert c.nsites(q) == 12 and c.nsites(params=p) == 4
    assert isinstance(c.version_string(), str) and c.version_string()


def test_cpp_symbols_exported():
    lib = os.path.join(PKG, "libftlm.so")
    if not os.path.exists(lib):
        pytest.skip("libftlm.so not built")
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    for sym in [
        "ftlm::nsites(ftlm::HubbardParams const&)",
        "ftlm::version_string()",
        "ftlm::build_hubbard_geometry(",
        "ftlm::FockBasis::FockBasis(int, int, int)",
        "ftlm::FockBasis::index_of(unsigned long, unsigned long) const",
        "ftlm::apply_extended_hubbard(",
        "ftlm::lanczos_tridiagonal(",
        "ftlm::lanczos_extrema(",
        "ftlm::tridiagonal_extrema(",
        "ftlm::ftlm_log_partition_complex(",
        "ftlm::ftlm_log_partition_real(",
        "ftlm::symmetry::OrbitKBlockBasis::build(",
        "ftlm::symmetry::OrbitKBlockBasis::map_raw_to_rep(",
        "ftlm::symmetry::OrbitKBlockBasis::reduced_index_of_rep(",
        "ftlm::symmetry::apply_orbit_k_block(",
        "ftlm::symmetry::fermionic_stabilizer_character_A_rect(",
    ]:
        assert sym in out, sym


def test_cpp_tests_built():
    bdir = os.path.join(ROOT, "tests", "cpp", "_bin")
    if not os.path.isdir(bdir):
        pytest.skip("C++ tests not built")
    for t in ("test_pipeline", "test_thermo", "test_orbit", "test_apply_host"):
        assert os.access(os.path.join(bdir, t), os.X_OK), t


# The reference's own symmetry tests, compiled against the drop-in headers and libftlm.so by
# oracle/Makefile (target reftests, where the reference checkout exists; the binaries live in
# oracle/_ref/tests).  These three use only host code and run here; test_momentum_partition_2x2
# and test_orbit_k_block_basis need the device (tests/test_gpu_cpp.py).
REFT = os.path.join(ROOT, "oracle", "_ref", "tests")
HOST_REF_TESTS = ("test_symmetry_translation_stage1", "test_symmetry_stage2", "test_symmetry_stage3")
ALL_REF_TESTS = HOST_REF_TESTS + ("test_momentum_partition_2x2", "test_orbit_k_block_basis")


def test_reference_symmetry_tests_built_against_dropin_headers():
    if not os.path.isdir("/root/reference/tests"):
        pytest.skip("no reference checkout here")
    for t in ALL_REF_TESTS:
        assert os.access(os.path.join(REFT, t), os.X_OK), f"{t}: run __graft_entry__.build()"


@pytest.mark.parametrize("prog", HOST_REF_TESTS)
def test_reference_symmetry_tests_pass_on_host(prog):
    import subprocess

    path = os.path.join(REFT, prog)
    if not os.path.exists(path):
        pytest.skip("reference tests not built (no reference checkout)")
    r = subprocess.run([path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{prog} ok" in r.stdout
