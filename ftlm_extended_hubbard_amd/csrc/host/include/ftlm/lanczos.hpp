// Drop-in for include/ftlm/lanczos.hpp.  The recurrence runs on the device.  An apply_h made by
// ftlm::hip::SectorOperator / KBlockOperator keeps the vectors in HBM; any other callable is
// called on host arrays once per step (the vector is copied out and back).
#pragma once

#include <complex>
#include <cstddef>
#include <functional>
#include <utility>
#include <vector>

namespace ftlm {

// The reference's three-vector host workspace.  The device path keeps its own vectors in HBM, so
// an instance passed in is accepted and left untouched.
struct LanczosComplexWorkspace {
  std::vector<std::complex<double>> q{};
  std::vector<std::complex<double>> q_prev{};
  std::vector<std::complex<double>> w{};
  void ensure(int dim) {
    const auto d = static_cast<std::size_t>(dim);
    q.resize(d);
    q_prev.assign(d, std::complex<double>(0.0, 0.0));
    w.resize(d);
  }
  std::size_t bytes_capacity() const {
    return (q.capacity() + q_prev.capacity() + w.capacity()) * sizeof(std::complex<double>);
  }
  void shrink_to_fit() {
    q.shrink_to_fit();
    q_prev.shrink_to_fit();
    w.shrink_to_fit();
  }
};

struct LanczosExtrema {
  double min_eval = 0.0;
  double max_eval = 0.0;
  int steps_used = 0;
};

using ApplyFn = std::function<void(const std::complex<double>* v, std::complex<double>* Hv)>;

// src/lanczos.cpp:211-231
LanczosExtrema lanczos_extrema(int dim, const ApplyFn& apply_h, int max_steps, unsigned seed = 1);

// src/lanczos.cpp:125-209: alpha gets `used` entries, beta `used - 1`; returns used (0 on failure).
int lanczos_tridiagonal(int dim, const ApplyFn& apply_h, int max_steps, unsigned seed, std::vector<double>* alpha,
                        std::vector<double>* beta, LanczosComplexWorkspace* ws = nullptr);

// src/lanczos.cpp:106-123 (120-rotation capped Jacobi), evaluated on the device.
std::pair<double, double> tridiagonal_extrema(const std::vector<double>& alpha, const std::vector<double>& beta,
                                              int steps);

}  // namespace ftlm
