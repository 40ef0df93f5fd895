// Drop-in for include/ftlm/fock_basis.hpp.  The sector is enumerated on the device (lexicographic
// order, index = i_up * D_dn + i_dn, as src/fock_basis.cpp:7-42); the two per-spin string tables
// (D_up + D_dn masks) are mirrored on the host for up_mask / down_mask / index_of.
#pragma once

#include <cstdint>
#include <memory>
#include <unordered_map>
#include <vector>

struct ftlm_op;

namespace ftlm {

class FockBasis {
 public:
  FockBasis(int n_sites, int n_up, int n_down);

  int n_sites() const { return n_sites_; }
  int n_up() const { return n_up_; }
  int n_down() const { return n_down_; }
  int dim() const { return dim_; }
  // std::out_of_range for an index outside [0, dim), as the reference's vector::at
  uint64_t up_mask(int idx) const;
  uint64_t down_mask(int idx) const;
  // index of (up, down), or -1 outside the sector
  int index_of(uint64_t up, uint64_t down) const;

  // The device sector (no Hamiltonian terms) behind this basis.
  const ftlm_op* device_handle() const { return dev_.get(); }

 private:
  int n_sites_, n_up_, n_down_, dim_;
  std::vector<uint32_t> up_str_, dn_str_;
  std::unordered_map<uint32_t, int> up_rank_, dn_rank_;
  std::shared_ptr<ftlm_op> dev_;
};

}  // namespace ftlm
