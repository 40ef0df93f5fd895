// Drop-in for include/ftlm/ftlm_thermo.hpp.
#pragma once

#include <complex>
#include <functional>
#include <vector>

#include "ftlm/lanczos.hpp"

namespace ftlm {

// The reference's host Jacobi scratch; the device quadrature keeps its own, so it is unused.
struct FtlmTridiagonalQuadratureScratch {
  std::vector<double> T_flat{};
  std::vector<double> V_flat{};
  void shrink_to_fit() {
    T_flat.shrink_to_fit();
    V_flat.shrink_to_fit();
  }
};

struct FtlmParams {
  int n_random = 16;
  int lanczos_steps = 96;
  unsigned seed = 1;
  LanczosComplexWorkspace* lanczos_ws = nullptr;
  FtlmTridiagonalQuadratureScratch* quad_scratch = nullptr;
};

// src/ftlm_thermo.cpp:150-211: H real symmetric, applied to the real and imaginary parts of the
// complex Lanczos vectors separately.
double ftlm_log_partition_real(int dim, const std::function<void(const double* x, double* y)>& apply_real, double beta,
                               const FtlmParams& par);

// src/ftlm_thermo.cpp:213-256.
double ftlm_log_partition_complex(int dim,
                                  const std::function<void(const std::complex<double>* x, std::complex<double>* y)>& apply_h,
                                  double beta, const FtlmParams& par);

}  // namespace ftlm
