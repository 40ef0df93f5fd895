// Drop-in for include/ftlm/symmetry/orbit.hpp: translation orbits of raw states.  An orbit is
// listed by the packs (up << 16) | dn of its distinct members in ascending order, and its
// canonical representative is the member of smallest pack.  The 4x4 forms are the rect forms at
// lx = ly = 4.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "ftlm/symmetry/packed_raw_state.hpp"
#include "ftlm/symmetry/raw_state.hpp"
#include "ftlm/symmetry/translation_group.hpp"

namespace ftlm {
namespace symmetry {

struct OrbitInfo {
  RawState representative{};
  std::uint16_t orbit_size = 0;
  std::uint8_t period_tx = 1;
  std::uint8_t period_ty = 1;
};

struct OrbitRecord {
  RawState representative{};
  std::uint16_t orbit_size = 0;
  std::uint8_t period_tx = 1;
  std::uint8_t period_ty = 1;
  std::vector<std::pair<std::int8_t, std::int8_t>> stabilizer;  // (dx, dy) fixing the representative
  std::uint16_t compatible_momentum_mask = 0;                    // bit ky * lx + kx
};

// Smallest p >= 1 with T_x^p s = s (at most lx), and the same along y.
inline std::uint8_t primitive_period_tx_rect(RawState s, int lx, int ly) {
  int p = 1;
  while (p < lx && translate_raw_state_rect(s, lx, ly, p, 0) != s) ++p;
  return static_cast<std::uint8_t>(p);
}
inline std::uint8_t primitive_period_ty_rect(RawState s, int lx, int ly) {
  int p = 1;
  while (p < ly && translate_raw_state_rect(s, lx, ly, 0, p) != s) ++p;
  return static_cast<std::uint8_t>(p);
}
inline std::uint8_t primitive_period_tx(RawState s) { return primitive_period_tx_rect(s, 4, 4); }
inline std::uint8_t primitive_period_ty(RawState s) { return primitive_period_ty_rect(s, 4, 4); }

// The distinct members of the orbit of seed, in ascending pack order.
inline void enumerate_translation_orbit_rect(RawState seed, int lx, int ly, std::vector<RawState>* out) {
  std::uint32_t packs[16 * 16];
  int n = 0;
  for_each_translation_rect(lx, ly,
                            [&](int dx, int dy) { packs[n++] = pack_raw_state(translate_raw_state_rect(seed, lx, ly, dx, dy)); });
  std::sort(packs, packs + n);
  out->clear();
  for (int i = 0; i < n; ++i)
    if (i == 0 || packs[i] != packs[i - 1]) out->push_back(unpack_raw_state(packs[i]));
}
inline void enumerate_translation_orbit(RawState seed, std::vector<RawState>* out) {
  enumerate_translation_orbit_rect(seed, 4, 4, out);
}

inline RawState canonical_raw_state_rect(RawState seed, int lx, int ly) {
  std::uint32_t best = pack_raw_state(seed);
  for_each_translation_rect(lx, ly, [&](int dx, int dy) {
    best = std::min(best, pack_raw_state(translate_raw_state_rect(seed, lx, ly, dx, dy)));
  });
  return unpack_raw_state(best);
}
inline RawState canonical_raw_state(RawState seed) { return canonical_raw_state_rect(seed, 4, 4); }

// Translations (dx, dy) with T s = s, dy outer and dx inner.
inline std::vector<std::pair<std::int8_t, std::int8_t>> translation_stabilizer_rect(RawState s, int lx, int ly) {
  std::vector<std::pair<std::int8_t, std::int8_t>> stab;
  for_each_translation_rect(lx, ly, [&](int dx, int dy) {
    if (translate_raw_state_rect(s, lx, ly, dx, dy) == s)
      stab.emplace_back(static_cast<std::int8_t>(dx), static_cast<std::int8_t>(dy));
  });
  return stab;
}
inline std::vector<std::pair<std::int8_t, std::int8_t>> translation_stabilizer(RawState s) {
  return translation_stabilizer_rect(s, 4, 4);
}

inline OrbitInfo make_orbit_info(RawState seed) {
  OrbitInfo info;
  info.representative = canonical_raw_state(seed);
  std::vector<RawState> members;
  enumerate_translation_orbit(info.representative, &members);
  info.orbit_size = static_cast<std::uint16_t>(members.size());
  info.period_tx = primitive_period_tx(info.representative);
  info.period_ty = primitive_period_ty(info.representative);
  return info;
}

// |orbit| * |stabiliser| = |Z4 x Z4|.
inline bool check_orbit_stabilizer_count_identity(const OrbitRecord& r) {
  return static_cast<int>(r.orbit_size) * static_cast<int>(r.stabilizer.size()) == translation_group_order();
}

}  // namespace symmetry
}  // namespace ftlm
