// Drop-in for include/ftlm/symmetry/momentum_sector.hpp: lattice momenta K = 2 pi (kx / lx,
// ky / ly), their Bloch phases and compatibility with an orbit's stabiliser, and the momentum
// sector map of a set of states (each translation orbit with the momenta it admits).
#pragma once

#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ftlm/symmetry/lattice_4x4.hpp"
#include "ftlm/symmetry/orbit.hpp"
#include "ftlm/symmetry/packed_raw_state.hpp"
#include "ftlm/symmetry/translation_group.hpp"

namespace ftlm {
namespace symmetry {

struct MomentumSector {
  int kx = 0;
  int ky = 0;
  int lx = 4;
  int ly = 4;
  constexpr int flat_index() const noexcept { return ky * lx + kx; }
  bool is_compatible(const OrbitRecord& r) const noexcept { return (r.compatible_momentum_mask >> flat_index()) & 1u; }
};

// exp(2 pi i (kx dx / lx + ky dy / ly)) == 1, in integers: kx dx ly + ky dy lx divisible by lx ly.
inline bool momentum_character_trivial_rect(int kx, int ky, int lx, int ly, int dx, int dy) {
  const long long v = static_cast<long long>(kx) * dx * ly + static_cast<long long>(ky) * dy * lx;
  return v % (static_cast<long long>(lx) * ly) == 0;
}
inline bool momentum_character_trivial(int kx, int ky, int dx, int dy) {
  return momentum_character_trivial_rect(kx, ky, 4, 4, dx, dy);
}

// exp(-2 pi i (kx dx / lx + ky dy / ly)) (momentum_sector.hpp:57-63).
inline std::complex<double> translation_bloch_phase(MomentumSector K, int dx, int dy) {
  const double th = -2.0 * std::acos(-1.0) *
                    (static_cast<double>(K.kx * dx) / static_cast<double>(K.lx) +
                     static_cast<double>(K.ky * dy) / static_cast<double>(K.ly));
  return {std::cos(th), std::sin(th)};
}

// 1 / sqrt(|G| |Stab|): the norm of an orbit's Bloch sum over the whole group.
inline double momentum_orbit_normalization_factor_rect(std::size_t stabilizer_size, int lx, int ly) noexcept {
  return 1.0 / std::sqrt(static_cast<double>(translation_group_order(lx, ly)) * static_cast<double>(stabilizer_size));
}
inline double momentum_orbit_normalization_factor(std::size_t stabilizer_size) noexcept {
  return momentum_orbit_normalization_factor_rect(stabilizer_size, 4, 4);
}

// Bit ky * lx + kx set where every stabiliser element has a trivial character at (kx, ky).
inline std::uint16_t compatible_momentum_mask_for_stabilizer(
    const std::vector<std::pair<std::int8_t, std::int8_t>>& stabilizer, int lx = 4, int ly = 4) {
  std::uint32_t mask = 0;
  for (int k = 0; k < lx * ly; ++k) {
    bool all = true;
    for (const auto& e : stabilizer) all = all && momentum_character_trivial_rect(k % lx, k / lx, lx, ly, e.first, e.second);
    if (all) mask |= 1u << k;
  }
  return static_cast<std::uint16_t>(mask);
}
inline void attach_compatible_momenta(OrbitRecord* r) {
  r->compatible_momentum_mask = compatible_momentum_mask_for_stabilizer(r->stabilizer, 4, 4);
}

struct MomentumSectorMap {
  int lx = 4;
  int ly = 4;
  std::vector<OrbitRecord> orbits;                                   // in order of first appearance
  std::vector<std::vector<std::size_t>> orbit_indices_by_momentum{};  // [ky * lx + kx] -> orbits
};

inline std::vector<RawState> unique_sorted_raw_states(std::vector<RawState> states) {
  std::sort(states.begin(), states.end(), RawStateLess{});
  states.erase(std::unique(states.begin(), states.end()), states.end());
  return states;
}

// The orbits of a set of states: the states are visited in ascending pack order and each orbit is
// recorded when its first member comes up, with its representative, size, periods, stabiliser and
// admitted momenta; then each momentum lists the orbits that admit it, in orbit order.
inline MomentumSectorMap build_momentum_sector_map_rect(std::vector<RawState> universe, int lx, int ly) {
  MomentumSectorMap map;
  map.lx = lx;
  map.ly = ly;
  map.orbit_indices_by_momentum.resize(static_cast<std::size_t>(lx * ly));
  std::unordered_map<std::uint32_t, std::size_t> seen;  // representative pack -> orbit
  for (RawState s : unique_sorted_raw_states(std::move(universe))) {
    const RawState rep = canonical_raw_state_rect(s, lx, ly);
    if (!seen.emplace(pack_raw_state(rep), map.orbits.size()).second) continue;
    OrbitRecord rec;
    rec.representative = rep;
    rec.stabilizer = translation_stabilizer_rect(rep, lx, ly);
    rec.orbit_size = static_cast<std::uint16_t>(translation_group_order(lx, ly) / static_cast<int>(rec.stabilizer.size()));
    rec.period_tx = primitive_period_tx_rect(rep, lx, ly);
    rec.period_ty = primitive_period_ty_rect(rep, lx, ly);
    rec.compatible_momentum_mask = compatible_momentum_mask_for_stabilizer(rec.stabilizer, lx, ly);
    map.orbits.push_back(std::move(rec));
  }
  for (std::size_t oi = 0; oi < map.orbits.size(); ++oi)
    for (int k = 0; k < lx * ly; ++k)
      if ((map.orbits[oi].compatible_momentum_mask >> k) & 1u) map.orbit_indices_by_momentum[static_cast<std::size_t>(k)].push_back(oi);
  return map;
}
inline MomentumSectorMap build_momentum_sector_map(std::vector<RawState> universe) {
  return build_momentum_sector_map_rect(std::move(universe), 4, 4);
}

inline int popcount_momentum_mask(std::uint16_t m) { return __builtin_popcount(m); }

}  // namespace symmetry
}  // namespace ftlm
