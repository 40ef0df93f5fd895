// Drop-in for include/ftlm/lattice.hpp: directed spinful hops, NN pairs and the rectangle.
#pragma once

#include <complex>
#include <vector>

#include "ftlm/hubbard_params.hpp"

namespace ftlm {

// H += coeff * c^dagger_{to, spin} c_{from, spin}; spin 0 = up, 1 = down.
struct SpinfulHopping {
  int from = 0;
  int to = 0;
  int spin = 0;
  std::complex<double> coeff{0.0, 0.0};
};

// Unordered nearest-neighbour pair (i < j) of the V n_i n_j term.
struct NearestPair {
  int i = 0;
  int j = 0;
};

struct RectLattice {
  int Lx = 1;
  int Ly = 1;
  int n_sites() const { return Lx * Ly; }
  int site_index(int x, int y) const { return y * Lx + x; }
  void coords(int site, int* x, int* y) const {
    *x = site % Lx;
    *y = site / Lx;
  }
};

// Same entries in the same order as the reference (src/lattice.cpp:38-110), including the
// duplicate bonds of short periodic directions; built by the device library.
void build_hubbard_geometry(const HubbardParams& p, const RectLattice& lat, std::vector<SpinfulHopping>* hoppings,
                            std::vector<NearestPair>* nn_pairs);

}  // namespace ftlm
