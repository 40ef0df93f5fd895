// Drop-in for include/ftlm/symmetry/fermionic_translation.hpp:12-54: the fermionic sign of a
// translation, the parity of the permutation that sorts the translated mode list back into the
// mode order (sites ascending, up (2s) before down (2s + 1) on each site).
#pragma once

#include <cstdint>

#include "ftlm/symmetry/raw_state.hpp"

namespace ftlm {
namespace symmetry {

inline int translated_site_rect(int site, int lx, int ly, int dx, int dy) {
  return ((site / lx + dy) % ly + ly) % ly * lx + ((site % lx + dx) % lx + lx) % lx;
}

// +1 / -1.  The modes are listed in their original order (site ascending, up before down) at
// their translated positions; each mode adds the number of earlier-listed modes that land above
// it, and the sign is the parity of that inversion count.
inline double fermionic_translation_sign_rect(RawState s, int lx, int ly, int dx, int dy) {
  std::uint64_t placed = 0;  // translated mode positions listed so far
  int inversions = 0;
  for (int site = 0; site < lx * ly; ++site) {
    const int t = translated_site_rect(site, lx, ly, dx, dy);
    for (int spin = 0; spin < 2; ++spin) {
      if (!(((spin ? s.dn : s.up) >> site) & 1u)) continue;
      const int m = 2 * t + spin;
      inversions += __builtin_popcountll(placed >> (m + 1));
      placed |= std::uint64_t{1} << m;
    }
  }
  return (inversions & 1) ? -1.0 : 1.0;
}

}  // namespace symmetry
}  // namespace ftlm
