// Drop-in for include/ftlm/hubbard_params.hpp: model parameters of the rectangular extended
// Hubbard lattice (same members and defaults).
#pragma once

namespace ftlm {

struct HubbardParams {
  int Lx = 2;
  int Ly = 2;
  double t = 1.0;
  double tp = 0.0;
  double U = 0.0;
  double V = 0.0;
  double phi_x = 0.0;
  double phi_y = 0.0;
};

// Lx * Ly (src/hubbard_params.cpp:5).
int nsites(const HubbardParams& p);

}  // namespace ftlm
