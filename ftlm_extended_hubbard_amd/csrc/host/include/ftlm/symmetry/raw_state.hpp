// Drop-in for include/ftlm/symmetry/raw_state.hpp: a configuration of at most 16 sites (one
// occupation pattern per spin) and its lattice translations.
#pragma once

#include <cstdint>

#include "ftlm/symmetry/occup_bits_translate.hpp"

namespace ftlm {
namespace symmetry {

struct RawState {
  std::uint16_t up = 0;
  std::uint16_t dn = 0;
};

inline bool operator==(RawState a, RawState b) { return a.up == b.up && a.dn == b.dn; }
inline bool operator!=(RawState a, RawState b) { return !(a == b); }

inline RawState translate_raw_state_rect(RawState s, int lx, int ly, int dx, int dy) {
  return RawState{translate_occupation_bits_rect(s.up, lx, ly, dx, dy),
                  translate_occupation_bits_rect(s.dn, lx, ly, dx, dy)};
}
inline RawState translate_raw_state(RawState s, int dx, int dy) { return translate_raw_state_rect(s, 4, 4, dx, dy); }
inline RawState translate_raw_state_Tx(RawState s) { return translate_raw_state(s, 1, 0); }
inline RawState translate_raw_state_Ty(RawState s) { return translate_raw_state(s, 0, 1); }

}  // namespace symmetry
}  // namespace ftlm
