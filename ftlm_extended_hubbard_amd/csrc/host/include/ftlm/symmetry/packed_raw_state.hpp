// Drop-in for include/ftlm/symmetry/packed_raw_state.hpp: (up << 16) | dn.
#pragma once

#include <cstdint>
#include <functional>

#include "ftlm/symmetry/raw_state.hpp"

namespace ftlm {
namespace symmetry {

inline constexpr std::uint32_t pack_raw_state(RawState s) noexcept {
  return (static_cast<std::uint32_t>(s.up) << 16) | static_cast<std::uint32_t>(s.dn);
}

inline constexpr RawState unpack_raw_state(std::uint32_t p) noexcept {
  return RawState{static_cast<std::uint16_t>(p >> 16), static_cast<std::uint16_t>(p & 0xFFFFu)};
}

struct PackedRawStateHash {
  std::size_t operator()(RawState s) const noexcept { return std::hash<std::uint32_t>{}(pack_raw_state(s)); }
};

struct RawStateLess {
  bool operator()(RawState a, RawState b) const noexcept { return pack_raw_state(a) < pack_raw_state(b); }
};

}  // namespace symmetry
}  // namespace ftlm
