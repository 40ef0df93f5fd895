// Drop-in C++ API: apply_extended_hubbard with host vectors (src/hubbard_hamiltonian.cpp:51-108),
// the call an unchanged reference driver makes in its hot loop (benchmarks/bench_2x2.cpp:405-435
// makes 200,000 of them).  Checks that repeated calls give identical results and that the
// operator is Hermitian (<x|H z> = <H x|z>), then prints the per-call wall time at dim 36 (2x2
// (2,2)), 4x3 (6,6) (dim 853,776) and, when the first argument is "big", 4x4 (8,8)
// (dim 165,636,900).

#include <chrono>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "check.hpp"
#include "ftlm/fock_basis.hpp"
#include "ftlm/hubbard_hamiltonian.hpp"
#include "ftlm/lattice.hpp"

using cplx = std::complex<double>;

namespace {

cplx dot(const std::vector<cplx>& a, const std::vector<cplx>& b) {
  cplx s(0.0, 0.0);
  for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
  return s;
}

void run(int Lx, int Ly, int nu, int nd, int calls, double phi) {
  ftlm::HubbardParams p;
  p.Lx = Lx;
  p.Ly = Ly;
  p.t = 1.0;
  p.tp = -0.25;
  p.U = 8.0;
  p.V = 1.0;
  p.phi_x = phi;
  ftlm::RectLattice lat;
  lat.Lx = Lx;
  lat.Ly = Ly;
  std::vector<ftlm::SpinfulHopping> hops;
  std::vector<ftlm::NearestPair> pairs;
  ftlm::build_hubbard_geometry(p, lat, &hops, &pairs);
  const ftlm::FockBasis b(Lx * Ly, nu, nd);
  const size_t n = static_cast<size_t>(b.dim());
  std::vector<cplx> x(n), z(n), hx(n), hz(n), hx2(n);
  std::mt19937 rng(5);
  std::normal_distribution<double> g(0.0, 1.0);
  for (size_t i = 0; i < n; ++i) {
    x[i] = cplx(g(rng), g(rng));
    z[i] = cplx(g(rng), g(rng));
  }
  ftlm::apply_extended_hubbard(p, b, hops, pairs, x.data(), hx.data());  // first call: builds and stages
  ftlm::apply_extended_hubbard(p, b, hops, pairs, z.data(), hz.data());
  const cplx a = dot(x, hz), c = dot(hx, z);
  CHECK(std::abs(a - c) <= 1e-10 * std::abs(a));
  const auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < calls; ++k) ftlm::apply_extended_hubbard(p, b, hops, pairs, x.data(), hx2.data());
  const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / calls;
  CHECK(std::memcmp(hx.data(), hx2.data(), n * sizeof(cplx)) == 0);
  std::printf("apply_extended_hubbard %dx%d (%d,%d) dim %zu: %.1f us per call (%d calls)\n", Lx, Ly, nu, nd, n, us,
              calls);
}

}  // namespace

int main(int argc, char** argv) {
  run(2, 2, 2, 2, 20000, 0.0);
  run(4, 3, 6, 6, 50, 0.05);
  if (argc > 1 && std::strcmp(argv[1], "big") == 0) run(4, 4, 8, 8, 5, 0.1);
  std::printf("test_apply_host ok\n");
  return 0;
}
