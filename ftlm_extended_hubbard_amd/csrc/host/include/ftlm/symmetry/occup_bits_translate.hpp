// Drop-in for include/ftlm/symmetry/occup_bits_translate.hpp: a 16-bit occupation pattern moved by
// a lattice translation (a particle on site s moves to site s translated by (dx, dy)).
#pragma once

#include <cstdint>

#include "ftlm/symmetry/lattice_4x4.hpp"

namespace ftlm {
namespace symmetry {

inline std::uint16_t translate_occupation_bits_rect(std::uint16_t bits, int lx, int ly, int dx, int dy) {
  const int n = lx * ly;
  const int sx = ((dx % lx) + lx) % lx, sy = ((dy % ly) + ly) % ly;
  std::uint32_t rest = bits & ((n >= 16) ? 0xFFFFu : ((1u << n) - 1u)), out = 0;
  while (rest) {
    const int s = __builtin_ctz(rest);
    rest &= rest - 1u;
    const int x = s % lx + sx, y = s / lx + sy;
    out |= 1u << ((y >= ly ? y - ly : y) * lx + (x >= lx ? x - lx : x));
  }
  return static_cast<std::uint16_t>(out);
}

inline std::uint16_t translate_occupation_bits(std::uint16_t bits, int dx, int dy) {
  return translate_occupation_bits_rect(bits, Lattice4x4::Lx, Lattice4x4::Ly, dx, dy);
}
inline std::uint16_t translate_occupation_Tx(std::uint16_t bits) { return translate_occupation_bits(bits, 1, 0); }
inline std::uint16_t translate_occupation_Ty(std::uint16_t bits) { return translate_occupation_bits(bits, 0, 1); }

}  // namespace symmetry
}  // namespace ftlm
