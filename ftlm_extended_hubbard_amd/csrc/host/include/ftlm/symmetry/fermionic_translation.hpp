// Drop-in for fermionic_translation_sign_rect (include/ftlm/symmetry/fermionic_translation.hpp:32-54).
#pragma once

#include "ftlm/symmetry/raw_state.hpp"

namespace ftlm {
namespace symmetry {

// +-1: parity of the reordering of the translated mode list (sites ascending, up before down).
double fermionic_translation_sign_rect(RawState s, int lx, int ly, int dx, int dy);

}  // namespace symmetry
}  // namespace ftlm
