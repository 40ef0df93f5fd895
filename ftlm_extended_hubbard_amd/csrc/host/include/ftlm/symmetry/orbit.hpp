// Drop-in for the orbit helpers of include/ftlm/symmetry/orbit.hpp:93-155.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "ftlm/symmetry/raw_state.hpp"

namespace ftlm {
namespace symmetry {

// Orbit member with the smallest pack.
RawState canonical_raw_state_rect(RawState s, int lx, int ly);

// Translations (dx, dy) fixing s, dy outer and dx inner.
std::vector<std::pair<std::int8_t, std::int8_t>> translation_stabilizer_rect(RawState s, int lx, int ly);

}  // namespace symmetry
}  // namespace ftlm
