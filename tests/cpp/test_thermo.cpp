// Drop-in C++ API, FTLM: the reference's tests/test_ftlm_thermo.cpp check (diag(2, 3), beta = 1,
// R = 48, M = 32, seed 12345 within 0.15 of the exact lnZ), the reference's own estimate for the
// same call (a fixture from the reference build), and the device-operator path on a sector.

#include <cmath>
#include <complex>
#include <vector>

#include "check.hpp"
#include "ftlm/ftlm_thermo.hpp"
#include "ftlm/hip/device.hpp"
#include "ftlm/hubbard_hamiltonian.hpp"
#include "ftlm/lattice.hpp"

int main() {
  {
    auto diag23 = [](const double* x, double* y) {
      y[0] = 2.0 * x[0];
      y[1] = 3.0 * x[1];
    };
    ftlm::FtlmParams par;
    par.n_random = 48;
    par.lanczos_steps = 32;
    par.seed = 12345u;
    const double exact = std::log(std::exp(-2.0) + std::exp(-3.0));
    const double est = ftlm::ftlm_log_partition_real(2, diag23, 1.0, par);
    CHECK_NEAR(est, exact, 0.15);
    // reference build's estimate for this call (tests/golden/golden.npz, key ftlm/diag23)
    CHECK_NEAR(est, -1.6757135346184655, 1e-9);
  }
  // dim <= 0 and dim == 1 follow the reference
  {
    ftlm::FtlmParams par;
    CHECK(std::isinf(ftlm::ftlm_log_partition_complex(0, [](const std::complex<double>*, std::complex<double>*) {},
                                                      1.0, par)));
    auto one = [](const std::complex<double>* x, std::complex<double>* y) { y[0] = 2.5 * x[0]; };
    CHECK_NEAR(ftlm::ftlm_log_partition_complex(1, one, 2.0, par), -5.0, 1e-15);
  }
  // a 4x2 sector: host-callback path and device-resident path give the same lnZ
  {
    ftlm::HubbardParams p;
    p.Lx = 4;
    p.Ly = 2;
    p.t = 1.0;
    p.tp = -0.25;
    p.U = 8.0;
    p.V = 1.0;
    ftlm::RectLattice lat{4, 2};
    std::vector<ftlm::SpinfulHopping> h;
    std::vector<ftlm::NearestPair> pr;
    ftlm::build_hubbard_geometry(p, lat, &h, &pr);
    ftlm::FockBasis b(8, 3, 2);
    ftlm::FtlmParams par;
    par.n_random = 2;
    par.lanczos_steps = 16;
    par.seed = 7;
    auto apply = [&](const std::complex<double>* x, std::complex<double>* y) {
      ftlm::apply_extended_hubbard(p, b, h, pr, x, y);
    };
    const double a = ftlm::ftlm_log_partition_complex(b.dim(), apply, 1.0, par);
    ftlm::hip::SectorOperator op(p, b, h, pr);
    const double c = ftlm::ftlm_log_partition_complex(op.dim(), op.apply_fn(), 1.0, par);
    CHECK_NEAR(a, c, 1e-10);
    const double r = ftlm::ftlm_log_partition_real(op.dim(), op.apply_real_fn(), 1.0, par);
    CHECK_NEAR(r, c, 1e-10);  // H is real here
  }
  std::cout << "test_thermo ok" << std::endl;
  return 0;
}
