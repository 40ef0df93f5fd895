// Drop-in C++ API, orbit k-blocks: the reference's tests/test_orbit_k_block_basis.cpp checks
// (|A| = |Stab| for surviving orbits, self-map of representatives, map count = sum of surviving
// orbit sizes, a 3x2 block builds) plus the reference's dimensions 12 and 66 and Hermiticity of
// the block apply, and the Gram blocks (HubbardMomentumBlock counterpart) of two 3x2 sectors.

#include <complex>
#include <utility>
#include <vector>

#include "check.hpp"
#include "ftlm/fock_basis.hpp"
#include "ftlm/lattice.hpp"
#include "ftlm/symmetry/momentum_sector.hpp"
#include "ftlm/symmetry/fermionic_translation.hpp"
#include "ftlm/symmetry/orbit.hpp"
#include "ftlm/symmetry/orbit_k_block_basis.hpp"
#include "ftlm/symmetry/packed_raw_state.hpp"
#include "ftlm/hip/device.hpp"

using namespace ftlm::symmetry;

int main() {
  {
    ftlm::FockBasis fb(4, 2, 2);
    MomentumSector K0{0, 0, 2, 2};
    OrbitKBlockBasis b0;
    CHECK(b0.build(fb, K0));
    CHECK(b0.dim() == 12);
    std::size_t sum_sizes = 0;
    for (const auto& e : b0.entries()) {
      sum_sizes += e.orbit_size;
      CHECK_NEAR(e.abs_A, static_cast<double>(e.stabilizer_size), 1e-9);
      CHECK(e.orbit_size * e.stabilizer_size == 4);
      CHECK(b0.reduced_index_of_rep(e.rep) >= 0);
      CHECK(canonical_raw_state_rect(e.rep, 2, 2) == e.rep);
      CHECK(translation_stabilizer_rect(e.rep, 2, 2).size() == e.stabilizer_size);
    }
    int n_ok = 0;
    for (int i = 0; i < fb.dim(); ++i) {
      const RawState s{static_cast<std::uint16_t>(fb.up_mask(i)), static_cast<std::uint16_t>(fb.down_mask(i))};
      int r = 0, dx = 0, dy = 0;
      double eta = 0.0;
      if (!b0.map_raw_to_rep(s, &r, &dx, &dy, &eta)) continue;
      ++n_ok;
      const RawState rep = b0.entries()[r].rep;
      CHECK(translate_raw_state_rect(rep, 2, 2, dx, dy) == s);
      CHECK_NEAR(eta, fermionic_translation_sign_rect(rep, 2, 2, dx, dy), 0.0);
      if (s == rep) {
        CHECK(dx == 0 && dy == 0);
        CHECK_NEAR(eta, 1.0, 0.0);
      }
    }
    CHECK(static_cast<std::size_t>(n_ok) == sum_sizes);
    CHECK(throws<std::invalid_argument>([] {
      ftlm::FockBasis big(18, 1, 1);
      OrbitKBlockBasis b;
      b.build(big, MomentumSector{0, 0, 6, 3});
    }));
  }
  {
    ftlm::FockBasis fb(6, 3, 3);
    OrbitKBlockBasis b;
    CHECK(b.build(fb, MomentumSector{1, 0, 3, 2}));
    CHECK(b.dim() == 66);
    // Hermiticity of H_K at zero flux on the 3x2 torus
    ftlm::HubbardParams p;
    p.Lx = 3;
    p.Ly = 2;
    p.t = 1.0;
    p.tp = -0.25;
    p.U = 8.0;
    p.V = 1.0;
    std::vector<ftlm::SpinfulHopping> h;
    std::vector<ftlm::NearestPair> pr;
    ftlm::build_hubbard_geometry(p, ftlm::RectLattice{3, 2}, &h, &pr);
    const int d = b.dim();
    std::vector<std::vector<std::complex<double>>> H(d, std::vector<std::complex<double>>(d));
    std::vector<std::complex<double>> e(d), col(d);
    for (int j = 0; j < d; ++j) {
      std::fill(e.begin(), e.end(), std::complex<double>(0.0, 0.0));
      e[j] = 1.0;
      apply_orbit_k_block(p, fb, b, h, pr, e.data(), col.data());
      for (int i = 0; i < d; ++i) H[i][j] = col[i];
    }
    double res = 0.0;
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) res = std::max(res, std::abs(H[i][j] - std::conj(H[j][i])));
    CHECK(res < 1e-12);
  }
  {
    // HubbardMomentumBlock counterpart: the Gram blocks of 3x2 (2,4) cover 216 of its 225 states
    // (the reference's k-gram run prints dim_k_total=216) and are Hermitian; a sector without
    // dropped orbits, (1,2), keeps all 90 states.
    ftlm::HubbardParams p;
    p.Lx = 3;
    p.Ly = 2;
    p.t = 1.0;
    p.tp = -0.35;
    p.U = 5.75;
    p.V = 0.9;
    std::vector<ftlm::SpinfulHopping> h;
    std::vector<ftlm::NearestPair> pr;
    ftlm::build_hubbard_geometry(p, ftlm::RectLattice{3, 2}, &h, &pr);
    for (const auto& sec : {std::pair<int, int>{2, 4}, std::pair<int, int>{1, 2}}) {
      ftlm::FockBasis fb(6, sec.first, sec.second);
      ftlm::hip::SectorOperator full(p, fb, h, pr);
      int total = 0;
      for (int ky = 0; ky < 2; ++ky)
        for (int kx = 0; kx < 3; ++kx) {
          const MomentumSector K{kx, ky, 3, 2};
          ftlm::hip::KBlockOperator kb(p, fb, K, h, pr);
          ftlm::hip::GramBlockOperator g(full, kb);
          const int d = g.dim();
          total += d;
          if (d == 0) continue;
          const auto apply = g.apply_fn();
          std::vector<std::vector<std::complex<double>>> H(d, std::vector<std::complex<double>>(d));
          std::vector<std::complex<double>> e(d), col(d);
          for (int j = 0; j < d; ++j) {
            std::fill(e.begin(), e.end(), std::complex<double>(0.0, 0.0));
            e[j] = 1.0;
            apply(e.data(), col.data());
            for (int i = 0; i < d; ++i) H[i][j] = col[i];
          }
          double res = 0.0;
          for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) res = std::max(res, std::abs(H[i][j] - std::conj(H[j][i])));
          CHECK(res < 1e-12);
        }
      CHECK(total == (sec.first == 2 ? 216 : 90));
    }
  }
  std::cout << "test_orbit ok" << std::endl;
  return 0;
}
