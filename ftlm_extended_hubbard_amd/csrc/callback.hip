// callback.hip -- operators whose H*v is a host function.
//
// The reference's Lanczos and FTLM entry points take the operator as
// std::function<void(const std::complex<double>* x, std::complex<double>* y)> over host arrays
// (include/ftlm/lanczos.hpp:47-63, include/ftlm/ftlm_thermo.hpp:41-50).  A callback operator lets
// any such function drive the device recurrence: the Krylov vectors, the FP64 reductions, the
// three-term update and the quadrature stay on the GPU; per H*v the current vector is copied to a
// pinned host buffer, the caller's function runs, and its result is copied back.

#include <algorithm>
#include <memory>

#include "common.hpp"

namespace ftlm_hip {

struct CallbackOp : ftlm_op {
  ftlm_host_apply_fn fn = nullptr;
  void* ctx = nullptr;
  DevBuf<double2> tmp;
  double2* hx = nullptr;
  double2* hy = nullptr;
  ~CallbackOp() override {
    if (hx) (void)hipHostFree(hx);
    if (hy) (void)hipHostFree(hy);
  }
};

namespace {

__global__ void k_gather_column(const double2* __restrict__ x, int R, int r, int64_t dim,
                                const double* __restrict__ xs, double2* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const double s = xs ? xs[r] : 1.0;
  const double2 v = x[i * R + r];
  out[i] = make_double2(v.x * s, v.y * s);
}

__global__ void k_scatter_column(const double2* __restrict__ in, int R, int r, int64_t dim,
                                 double2* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < dim) y[i * R + r] = in[i];
}

// Per-block partial sums of <xs_r x_r | y_r> for every vector r: dotp[block * R + r].
__global__ __launch_bounds__(256) void k_dot_partials(const double2* __restrict__ x, const double2* __restrict__ y,
                                                      int R, int64_t dim, const double* __restrict__ xs,
                                                      double2* __restrict__ dotp) {
  __shared__ double red[4][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int r = 0; r < R; ++r) {
    const double s = xs ? xs[r] : 1.0;
    double a = 0.0, b = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < dim;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
      const double2 q = x[i * R + r], w = y[i * R + r];
      const double qr = s * q.x, qi = s * q.y;
      a += qr * w.x + qi * w.y;
      b += qr * w.y - qi * w.x;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      a += __shfl_xor(a, o, 64);
      b += __shfl_xor(b, o, 64);
    }
    __syncthreads();
    if (lane == 0) {
      red[wv][0] = a;
      red[wv][1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      dotp[static_cast<int64_t>(blockIdx.x) * R + r] =
          make_double2(red[0][0] + red[1][0] + red[2][0] + red[3][0], red[0][1] + red[1][1] + red[2][1] + red[3][1]);
  }
}

}  // namespace

int hv_apply_callback(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                      hipStream_t st) {
  const CallbackOp* C = static_cast<const CallbackOp*>(op);
  FTLM_REQUIRE(dtype == FTLM_C128, FTLM_EUNSUPPORTED, "host-callback operators take complex128 vectors");
  const int64_t dim = C->dim;
  if (dim == 0) return 0;
  const double2* xp = static_cast<const double2*>(x);
  double2* yp = static_cast<double2*>(y);
  const size_t bytes = static_cast<size_t>(dim) * sizeof(double2);
  for (int r = 0; r < R; ++r) {
    hipLaunchKernelGGL(k_gather_column, dim3(grid_for(dim, 256)), dim3(256), 0, st, xp, R, r, dim, xs, C->tmp.p);
    check_launch("k_gather_column");
    FTLM_HIP_CHECK(hipMemcpyAsync(C->hx, C->tmp.p, bytes, hipMemcpyDeviceToHost, st));
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
    C->fn(reinterpret_cast<const double*>(C->hx), reinterpret_cast<double*>(C->hy), C->ctx);
    FTLM_HIP_CHECK(hipMemcpyAsync(C->tmp.p, C->hy, bytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scatter_column, dim3(grid_for(dim, 256)), dim3(256), 0, st, C->tmp.p, R, r, dim, yp);
    check_launch("k_scatter_column");
  }
  if (!dotp) return 0;
  const int grid = static_cast<int>(std::min<int64_t>(1024, (dim + 255) / 256));
  hipLaunchKernelGGL(k_dot_partials, dim3(grid), dim3(256), 0, st, xp, static_cast<const double2*>(yp), R, dim, xs,
                     dotp);
  check_launch("k_dot_partials");
  return grid;
}

}  // namespace ftlm_hip

using namespace ftlm_hip;

extern "C" {

int ftlm_hip_callback_op_create(int64_t dim, ftlm_host_apply_fn apply, void* ctx, ftlm_op_t** out) {
  return guarded([&] {
    FTLM_REQUIRE(out != nullptr && apply != nullptr, FTLM_EINVAL, "null argument");
    FTLM_REQUIRE(dim >= 0, FTLM_EINVAL, "negative dimension");
    *out = nullptr;
    int dev = 0;
    FTLM_HIP_CHECK(hipGetDevice(&dev));
    auto C = std::make_unique<CallbackOp>();
    C->kind = kCallback;
    C->device = dev;
    C->dim = dim;
    C->fn = apply;
    C->ctx = ctx;
    if (dim > 0) {
      C->tmp.alloc(static_cast<size_t>(dim));
      const size_t bytes = static_cast<size_t>(dim) * sizeof(double2);
      FTLM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&C->hx), bytes, hipHostMallocDefault));
      FTLM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&C->hy), bytes, hipHostMallocDefault));
    }
    *out = C.release();
  });
}

}  // extern "C"
