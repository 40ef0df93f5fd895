// Device-resident operators for the drop-in API.  An ApplyFn returned by apply_fn() is recognised
// by lanczos_tridiagonal / lanczos_extrema / ftlm_log_partition_complex, which then run on the
// operator without copying vectors to the host; called directly it computes y = H x on host
// arrays like any ApplyFn.
#pragma once

#include <complex>
#include <functional>
#include <memory>
#include <vector>

#include "ftlm/fock_basis.hpp"
#include "ftlm/hubbard_params.hpp"
#include "ftlm/lattice.hpp"
#include "ftlm/lanczos.hpp"
#include "ftlm/symmetry/orbit_k_block_basis.hpp"

struct ftlm_op;

namespace ftlm {
namespace hip {

// Callable stored in the std::function returned by apply_fn().
struct DeviceApply {
  std::shared_ptr<ftlm_op> op;
  void operator()(const std::complex<double>* x, std::complex<double>* y) const;
};

// Re(H) of a full-sector operator for ftlm_log_partition_real.
struct DeviceApplyReal {
  std::shared_ptr<ftlm_op> op;
  void operator()(const double* x, double* y) const;
};

class SectorOperator {
 public:
  SectorOperator(const HubbardParams& p, const FockBasis& basis, const std::vector<SpinfulHopping>& hoppings,
                 const std::vector<NearestPair>& nn_pairs);
  int dim() const { return dim_; }
  ApplyFn apply_fn() const { return DeviceApply{op_}; }
  std::function<void(const double*, double*)> apply_real_fn() const { return DeviceApplyReal{op_}; }
  const ftlm_op* handle() const { return op_.get(); }
  std::shared_ptr<ftlm_op> shared_handle() const { return op_; }

 private:
  int dim_ = 0;
  std::shared_ptr<ftlm_op> op_;
};

class KBlockOperator {
 public:
  // `share`: another block of the same sector and Hamiltonian whose orbit graph is reused.
  KBlockOperator(const HubbardParams& p, const FockBasis& fb, symmetry::MomentumSector K,
                 const std::vector<SpinfulHopping>& hoppings, const std::vector<NearestPair>& nn_pairs,
                 double survive_tol = 1e-12, const KBlockOperator* share = nullptr);
  int dim() const { return dim_; }
  ApplyFn apply_fn() const { return DeviceApply{op_}; }
  const ftlm_op* handle() const { return op_.get(); }

 private:
  int dim_ = 0;
  std::shared_ptr<ftlm_op> op_;
};

// HubbardMomentumBlock (include/ftlm/symmetry/hubbard_momentum_action.hpp:66-113): the momentum
// block of `sector` in the reference's Gram subspace, applied through the full sector on the
// device.  `kblock` (same sector and K) supplies the orbits; the block keeps the sector operator
// alive and does not need `kblock` afterwards.
class GramBlockOperator {
 public:
  GramBlockOperator(const SectorOperator& sector, const KBlockOperator& kblock);
  int dim() const { return dim_; }
  ApplyFn apply_fn() const { return DeviceApply{op_}; }
  const ftlm_op* handle() const { return op_.get(); }

 private:
  int dim_ = 0;
  std::shared_ptr<ftlm_op> sector_;
  std::shared_ptr<ftlm_op> op_;
};

}  // namespace hip
}  // namespace ftlm
