// Drop-in for include/ftlm/symmetry/k_block_matrix_free.hpp:22-39: the abstract matrix-free
// momentum-block operator and the adapter that exposes any block type with dim() and
// apply(x, y) (HubbardMomentumBlock in the reference; here GramBlockOperator or
// ftlm::hip::KBlockOperator) through it.
#pragma once

#include <complex>

namespace ftlm {
namespace symmetry {
namespace k_block_matrix_free {

struct KBlockHamiltonianApply {
  virtual ~KBlockHamiltonianApply() = default;
  virtual int dim() const = 0;
  virtual void apply(const std::complex<double>* x, std::complex<double>* y) const = 0;
};

template <typename Block>
class HubbardBlockAdapter final : public KBlockHamiltonianApply {
 public:
  explicit HubbardBlockAdapter(const Block* block) : block_(block) {}
  int dim() const override { return static_cast<int>(block_->dim()); }
  void apply(const std::complex<double>* x, std::complex<double>* y) const override { block_->apply(x, y); }

 private:
  const Block* block_ = nullptr;
};

}  // namespace k_block_matrix_free
}  // namespace symmetry
}  // namespace ftlm
