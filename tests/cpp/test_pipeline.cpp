// Drop-in C++ API, full sectors: the checks of the reference's tests/test_hubbard_pipeline.cpp
// (basis size and index, Lanczos ground state against a dense diagonalisation, Hermiticity
// under flux) plus the reference's error behaviour and the device-resident operator path.

#include <algorithm>
#include <complex>
#include <vector>

#include "check.hpp"
#include "ftlm/fock_basis.hpp"
#include "ftlm/hip/device.hpp"
#include "ftlm/hubbard_hamiltonian.hpp"
#include "ftlm/lanczos.hpp"
#include "ftlm/lattice.hpp"
#include "ftlm/version.hpp"

using cplx = std::complex<double>;

namespace {

std::vector<std::vector<cplx>> dense(const ftlm::HubbardParams& p, const ftlm::FockBasis& b,
                                     const std::vector<ftlm::SpinfulHopping>& h,
                                     const std::vector<ftlm::NearestPair>& pr) {
  const int n = b.dim();
  std::vector<std::vector<cplx>> H(n, std::vector<cplx>(n));
  std::vector<cplx> e(n), col(n);
  for (int j = 0; j < n; ++j) {
    std::fill(e.begin(), e.end(), cplx(0.0, 0.0));
    e[j] = 1.0;
    ftlm::apply_extended_hubbard(p, b, h, pr, e.data(), col.data());
    for (int i = 0; i < n; ++i) H[i][j] = col[i];
  }
  return H;
}

// Smallest eigenvalue of a real symmetric matrix by cyclic Jacobi sweeps run to convergence.
double min_eig(std::vector<std::vector<double>> A) {
  const int n = static_cast<int>(A.size());
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[i][j] * A[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (std::abs(A[p][q]) < 1e-300) continue;
        const double th = 0.5 * std::atan2(2.0 * A[p][q], A[q][q] - A[p][p]);
        const double c = std::cos(th), s = std::sin(th);
        for (int k = 0; k < n; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
      }
  }
  double lo = A[0][0];
  for (int i = 1; i < n; ++i) lo = std::min(lo, A[i][i]);
  return lo;
}

}  // namespace

int main() {
  CHECK(std::string(ftlm::version_string()).size() > 0);

  // basis size and index
  {
    ftlm::FockBasis b(2, 1, 1);
    CHECK(b.dim() == 4);
    CHECK(b.index_of(1ull << 0, 1ull << 1) >= 0);
    CHECK(b.index_of(3ull, 1ull) == -1);
    CHECK(throws<std::out_of_range>([&] { (void)b.up_mask(4); }));
    CHECK(throws<std::invalid_argument>([] { ftlm::FockBasis bad(4, 5, 0); }));
    CHECK(throws<std::invalid_argument>([] { ftlm::FockBasis bad(0, 0, 0); }));
    ftlm::FockBasis c(4, 2, 1);
    for (int i = 0; i < c.dim(); ++i) CHECK(c.index_of(c.up_mask(i), c.down_mask(i)) == i);
  }

  // 2x1 periodic chain, kinetic only: Lanczos ground state against the dense matrix
  {
    ftlm::HubbardParams p;
    p.Lx = 2;
    p.Ly = 1;
    p.t = 1.0;
    ftlm::RectLattice lat{2, 1};
    std::vector<ftlm::SpinfulHopping> h;
    std::vector<ftlm::NearestPair> pr;
    ftlm::build_hubbard_geometry(p, lat, &h, &pr);
    ftlm::FockBasis b(lat.n_sites(), 1, 1);
    const auto H = dense(p, b, h, pr);
    std::vector<std::vector<double>> Hr(b.dim(), std::vector<double>(b.dim()));
    for (int i = 0; i < b.dim(); ++i)
      for (int j = 0; j < b.dim(); ++j) {
        CHECK(std::abs(H[i][j].imag()) < 1e-12);
        CHECK(std::abs(H[i][j] - std::conj(H[j][i])) < 1e-12);
        Hr[i][j] = H[i][j].real();
      }
    const double e0 = min_eig(Hr);
    auto apply = [&](const cplx* v, cplx* w) { ftlm::apply_extended_hubbard(p, b, h, pr, v, w); };
    const ftlm::LanczosExtrema lz = ftlm::lanczos_extrema(b.dim(), apply, 48, 42);
    CHECK_NEAR(lz.min_eval, e0, 5e-5);
    // the device-resident operator gives the same recurrence
    ftlm::hip::SectorOperator op(p, b, h, pr);
    const ftlm::LanczosExtrema lz2 = ftlm::lanczos_extrema(op.dim(), op.apply_fn(), 48, 42);
    CHECK(lz2.steps_used == lz.steps_used);
    CHECK_NEAR(lz2.min_eval, lz.min_eval, 1e-10);
  }

  // flux on the 2x2 torus: H stays Hermitian
  {
    ftlm::HubbardParams p;
    p.Lx = 2;
    p.Ly = 2;
    p.t = 1.0;
    p.phi_x = 0.31;
    p.phi_y = -0.27;
    ftlm::RectLattice lat{2, 2};
    std::vector<ftlm::SpinfulHopping> h;
    std::vector<ftlm::NearestPair> pr;
    ftlm::build_hubbard_geometry(p, lat, &h, &pr);
    ftlm::FockBasis b(lat.n_sites(), 2, 2);
    const auto H = dense(p, b, h, pr);
    double res = 0.0;
    for (int i = 0; i < b.dim(); ++i)
      for (int j = 0; j < b.dim(); ++j) res = std::max(res, std::abs(H[i][j] - std::conj(H[j][i])));
    CHECK(res < 1e-10);
  }

  // tridiagonal extrema of a 2x2 tridiagonal [[1, 2], [2, 1]] -> (-1, 3)
  {
    const auto mm = ftlm::tridiagonal_extrema({1.0, 1.0}, {2.0}, 2);
    CHECK_NEAR(mm.first, -1.0, 1e-12);
    CHECK_NEAR(mm.second, 3.0, 1e-12);
  }

  std::cout << "test_pipeline ok" << std::endl;
  return 0;
}
