// Drop-in for include/ftlm/symmetry/k_basis.hpp: the canonical representatives that admit one
// momentum, sorted by pack, with an index lookup; and the member states of those orbits (the
// seeds of the Bloch columns of the Gram path).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <vector>

#include "ftlm/symmetry/momentum_sector.hpp"
#include "ftlm/symmetry/orbit.hpp"
#include "ftlm/symmetry/packed_raw_state.hpp"
#include "ftlm/symmetry/raw_state.hpp"

namespace ftlm {
namespace symmetry {

struct KBasis {
  static constexpr std::size_t npos = std::numeric_limits<std::size_t>::max();

  MomentumSector K{};
  std::vector<std::uint32_t> rep_packed{};   // ascending
  std::vector<RawState> representatives{};  // representatives[i] unpacks rep_packed[i]

  std::size_t dim() const noexcept { return representatives.size(); }

  static KBasis build(const MomentumSectorMap& map, MomentumSector Ksec) {
    KBasis b;
    b.K = Ksec;
    b.K.lx = map.lx;
    b.K.ly = map.ly;
    const std::size_t k = static_cast<std::size_t>(Ksec.flat_index());
    if (k < map.orbit_indices_by_momentum.size()) {
      for (std::size_t oi : map.orbit_indices_by_momentum[k]) b.rep_packed.push_back(pack_raw_state(map.orbits[oi].representative));
      std::sort(b.rep_packed.begin(), b.rep_packed.end());
      b.rep_packed.erase(std::unique(b.rep_packed.begin(), b.rep_packed.end()), b.rep_packed.end());
      for (std::uint32_t p : b.rep_packed) b.representatives.push_back(unpack_raw_state(p));
    }
    return b;
  }

  // Row of a canonical representative, or npos.
  std::size_t index_of(RawState rep) const noexcept {
    const std::uint32_t p = pack_raw_state(rep);
    const auto it = std::lower_bound(rep_packed.begin(), rep_packed.end(), p);
    return (it != rep_packed.end() && *it == p) ? static_cast<std::size_t>(it - rep_packed.begin()) : npos;
  }
};

// Every member of every orbit that admits K, ascending by pack and distinct.
inline std::vector<RawState> momentum_phi_seeds(const MomentumSectorMap& map, MomentumSector K) {
  std::vector<RawState> seeds;
  const std::size_t k = static_cast<std::size_t>(K.flat_index());
  if (k >= map.orbit_indices_by_momentum.size()) return seeds;
  std::vector<RawState> members;
  for (std::size_t oi : map.orbit_indices_by_momentum[k]) {
    enumerate_translation_orbit_rect(map.orbits[oi].representative, map.lx, map.ly, &members);
    seeds.insert(seeds.end(), members.begin(), members.end());
  }
  return unique_sorted_raw_states(std::move(seeds));
}

}  // namespace symmetry
}  // namespace ftlm
