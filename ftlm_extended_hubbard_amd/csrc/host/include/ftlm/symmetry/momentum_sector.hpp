// Drop-in for the momentum pieces of include/ftlm/symmetry/momentum_sector.hpp.
#pragma once

#include <complex>

namespace ftlm {
namespace symmetry {

struct MomentumSector {
  int kx = 0;
  int ky = 0;
  int lx = 4;
  int ly = 4;
  constexpr int flat_index() const noexcept { return ky * lx + kx; }
};

// exp(-2 pi i (kx dx / lx + ky dy / ly)) (momentum_sector.hpp:57-63)
std::complex<double> translation_bloch_phase(MomentumSector K, int dx, int dy);

}  // namespace symmetry
}  // namespace ftlm
