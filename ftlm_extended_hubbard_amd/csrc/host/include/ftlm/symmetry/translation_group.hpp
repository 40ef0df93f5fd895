// Drop-in for include/ftlm/symmetry/translation_group.hpp: the translation group Z_lx x Z_ly of a
// rectangular torus and its elements (dx, dy), visited dy outer, dx inner.
#pragma once

namespace ftlm {
namespace symmetry {

inline constexpr int translation_group_order(int lx, int ly) noexcept { return lx * ly; }
inline constexpr int translation_group_order() noexcept { return translation_group_order(4, 4); }

template <class F>
void for_each_translation_rect(int lx, int ly, F&& f) {
  for (int t = 0; t < lx * ly; ++t) f(t % lx, t / lx);
}

template <class F>
void for_each_translation_z4z4(F&& f) {
  for_each_translation_rect(4, 4, f);
}

}  // namespace symmetry
}  // namespace ftlm
