// Drop-in for include/ftlm/hubbard_hamiltonian.hpp.
#pragma once

#include <complex>
#include <vector>

#include "ftlm/fock_basis.hpp"
#include "ftlm/hubbard_params.hpp"
#include "ftlm/lattice.hpp"

namespace ftlm {

// y = H x on host arrays of length basis.dim() (y is overwritten, as src/hubbard_hamiltonian.cpp:56
// does).  The device operator for (basis, hoppings, nn_pairs, U, V) is built on first use and
// cached; every call copies x to the device and y back.  Vector-resident callers use
// ftlm::hip::SectorOperator (ftlm/hip/device.hpp) instead.
void apply_extended_hubbard(const HubbardParams& p, const FockBasis& basis, const std::vector<SpinfulHopping>& hoppings,
                            const std::vector<NearestPair>& nn_pairs, const std::complex<double>* x,
                            std::complex<double>* y);

}  // namespace ftlm
