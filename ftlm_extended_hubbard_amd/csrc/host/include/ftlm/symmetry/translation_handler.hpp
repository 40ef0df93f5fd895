// Drop-in for include/ftlm/symmetry/translation_handler.hpp: the 4x4 translation helpers of
// orbit.hpp behind one name.
#pragma once

#include <vector>

#include "ftlm/symmetry/orbit.hpp"
#include "ftlm/symmetry/raw_state.hpp"

namespace ftlm {
namespace symmetry {

struct TranslationHandler {
  static RawState canonical(RawState s) { return canonical_raw_state(s); }
  static void orbit(RawState seed, std::vector<RawState>* out) { enumerate_translation_orbit(seed, out); }
  static OrbitInfo info(RawState seed) { return make_orbit_info(seed); }
};

}  // namespace symmetry
}  // namespace ftlm
