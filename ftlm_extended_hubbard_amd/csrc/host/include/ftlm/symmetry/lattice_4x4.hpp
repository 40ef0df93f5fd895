// Drop-in for include/ftlm/symmetry/lattice_4x4.hpp: the 4x4 periodic torus, sites numbered
// row-major (site = y * Lx + x), with periodic translations of site indices.
#pragma once

namespace ftlm {
namespace symmetry {

struct Lattice4x4 {
  static constexpr int Lx = 4;
  static constexpr int Ly = 4;
  static constexpr int N = Lx * Ly;

  static constexpr int site_from_xy(int x, int y) { return y * Lx + x; }
  static constexpr void xy_from_site(int site, int* x, int* y) {
    *x = site % Lx;
    *y = site / Lx;
  }
  // a mod m in [0, m) for m > 0, negative a included
  static constexpr int imod(int a, int m) { return ((a % m) + m) % m; }
  static constexpr int translate_site_index(int site, int dx, int dy) {
    return site_from_xy(imod(site % Lx + dx, Lx), imod(site / Lx + dy, Ly));
  }
  static constexpr int translate_site_Tx(int site) { return translate_site_index(site, 1, 0); }
  static constexpr int translate_site_Ty(int site) { return translate_site_index(site, 0, 1); }
};

}  // namespace symmetry
}  // namespace ftlm
