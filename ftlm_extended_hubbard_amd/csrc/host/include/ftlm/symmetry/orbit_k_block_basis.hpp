// Drop-in for include/ftlm/symmetry/orbit_k_block_basis.hpp.  The orbit graph, A_K, the surviving
// rows and the block operator live on the device.
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "ftlm/fock_basis.hpp"
#include "ftlm/hubbard_params.hpp"
#include "ftlm/lattice.hpp"
#include "ftlm/symmetry/momentum_sector.hpp"
#include "ftlm/symmetry/raw_state.hpp"

struct ftlm_op;

namespace ftlm {
namespace symmetry {

// chi_K(R) = conj(translation_bloch_phase(K, dx, dy))
inline std::complex<double> momentum_character_chi_K(MomentumSector K, int dx, int dy) {
  return std::conj(translation_bloch_phase(K, dx, dy));
}

// A_alpha(K) = sum over the stabiliser list of eta(rep, S) * translation_bloch_phase(K, S).
std::complex<double> fermionic_stabilizer_character_A_rect(
    const RawState& rep, MomentumSector K, int lx, int ly,
    const std::vector<std::pair<std::int8_t, std::int8_t>>& stabilizer);

struct OrbitKRepEntry {
  RawState rep{};
  std::size_t orbit_index = 0;  // position of the orbit in this build's pack-sorted orbit list
  std::uint16_t orbit_size = 0;
  std::uint16_t stabilizer_size = 0;
  std::complex<double> A_k{0.0, 0.0};
  double abs_A = 0.0;
};

class OrbitKBlockBasis {
 public:
  MomentumSector K{};
  int lx = 0;
  int ly = 0;

  // std::invalid_argument for more than 16 sites, as the reference.
  bool build(const FockBasis& fb, MomentumSector K, double survive_tol = 1e-12);
  int dim() const noexcept { return dim_; }
  const std::vector<OrbitKRepEntry>& entries() const;
  double survive_tol() const noexcept { return survive_tol_; }
  int reduced_index_of_rep(RawState rep) const;
  bool map_raw_to_rep(RawState s, int* reduced_index, int* dx, int* dy, double* eta) const;

  // The device block (basis data only; apply_orbit_k_block adds the Hamiltonian).
  const ftlm_op* device_handle() const { return dev_.get(); }
  int n_up() const { return n_up_; }
  int n_down() const { return n_dn_; }

 private:
  int dim_ = 0, n_up_ = 0, n_dn_ = 0;
  double survive_tol_ = 1e-12;
  std::shared_ptr<ftlm_op> dev_;
  mutable std::vector<OrbitKRepEntry> entries_;
  mutable bool have_entries_ = false;
};

// y = H_K x (src/orbit_k_block_basis.cpp:151-220) on host arrays of length basis.dim(); the
// device block with the Hamiltonian is built on first use and cached.
void apply_orbit_k_block(const HubbardParams& p, const FockBasis& fb, const OrbitKBlockBasis& basis,
                         const std::vector<SpinfulHopping>& hoppings, const std::vector<NearestPair>& nn_pairs,
                         const std::complex<double>* x, std::complex<double>* y);

}  // namespace symmetry
}  // namespace ftlm
