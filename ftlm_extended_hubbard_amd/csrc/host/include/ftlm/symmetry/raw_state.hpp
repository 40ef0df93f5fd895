// Drop-in for include/ftlm/symmetry/raw_state.hpp: a configuration of at most 16 sites.
#pragma once

#include <cstdint>

namespace ftlm {
namespace symmetry {

struct RawState {
  std::uint16_t up = 0;
  std::uint16_t dn = 0;
};

inline bool operator==(RawState a, RawState b) { return a.up == b.up && a.dn == b.dn; }
inline bool operator!=(RawState a, RawState b) { return !(a == b); }

// Both strings translated by (dx, dy) on the lx x ly torus (device).
RawState translate_raw_state_rect(RawState s, int lx, int ly, int dx, int dy);

}  // namespace symmetry
}  // namespace ftlm
