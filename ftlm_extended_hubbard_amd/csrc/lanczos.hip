// lanczos.hip -- device-resident block Lanczos recurrences and the FTLM drivers.
//
// Replaces lanczos_tridiagonal (src/lanczos.cpp:125-209), lanczos_extrema (src/lanczos.cpp:211-231)
// and the sample loop of ftlm_log_partition_complex / _real (src/ftlm_thermo.cpp:150-256).
//
// R independent recurrences (one per random start vector) advance together as a block so that
// each H*V launch shares its decode work across the R vectors.  Per step and vector r:
//     w = H q;  z = <q|w>;  alpha = Re z;  w -= z q;  w -= beta_prev q_prev;
//     if |w|^2 < 1e-24: stop r;  beta = |w|;  q_prev = q;  q = w / beta
// exactly as the reference (no reorthogonalisation; complex z is subtracted).  Dots and norms
// are FP64 (per-block partials, then a fixed-order reduction, so results are deterministic).
// Normalisation is lazy: buffer A holds w and a per-vector scale sa = 1/beta, applied inside the
// H*V and update kernels, which saves one read+write of the block per step.  Three buffers
// rotate (A, B, C) -> (C, A, B); the scalar bookkeeping (alpha, beta, scales, stop flags) stays
// on the device, so the whole step sequence can be recorded into a hipGraph.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <thread>
#include <vector>

#include "common.hpp"
#include "gauss_mt.hpp"

namespace ftlm_hip {

void quadrature_device(const double* d_alpha, const double* d_beta, const int* d_n, const int* d_nb, int stride,
                       int nsamp, int max_n, int quad, int cap_fixed, double tol, int with_weights, double* d_ritz,
                       double* d_weight, int out_stride, hipStream_t st);
void thermo_device(const double* d_ritz, const double* d_weight, const int* d_used, int n_ops, int R, int M,
                   const int64_t* d_dims, const int* d_N, const double* d_betas, int nb, const double* d_mus, int nmu,
                   double* d_lnZ, double* d_E, double* d_lnZgc, double* d_Ngc, double* d_Egc, hipStream_t st);

template <typename T>
struct V2Of;
template <>
struct V2Of<double> {
  using t = double2;
};
template <>
struct V2Of<float> {
  using t = float2;
};

constexpr int kRedBlock = 256;
constexpr int kUpdBlock = 256;
constexpr int kUpdMaxGrid = 4096;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block r: z_r = sum_b dotp[b][r]; alpha[k][r] = Re z_r for live vectors.
__global__ __launch_bounds__(kRedBlock) void k_reduce_alpha(const double2* __restrict__ dotp, int nblk, int R,
                                                            const int* __restrict__ alive, double2* __restrict__ z,
                                                            double* __restrict__ alpha, int k) {
  const int r = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) {
    const double2 v = dotp[static_cast<int64_t>(i) * R + r];
    a += v.x;
    b += v.y;
  }
  __shared__ double sa[kRedBlock / 64], sb[kRedBlock / 64];
  a = wsum(a);
  b = wsum(b);
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int w = 0; w < kRedBlock / 64; ++w) {
      x += sa[w];
      y += sb[w];
    }
    if (alive[r]) {
      z[r] = make_double2(x, y);
      alpha[static_cast<int64_t>(k) * R + r] = x;
    } else {
      z[r] = make_double2(0.0, 0.0);
    }
  }
}

// Flat form of the update over the n = dim*R complex values of the block: consecutive threads
// touch consecutive values (fully coalesced).  The block size is a multiple of R and so is the
// grid stride, so each thread always sees the same vector r = tid % R and keeps one norm partial.
template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_update_flat(int64_t n, int R, typename V2Of<T>::t* __restrict__ C,
                                                           const typename V2Of<T>::t* __restrict__ A,
                                                           const typename V2Of<T>::t* __restrict__ B,
                                                           const double2* __restrict__ z, const double* __restrict__ sa,
                                                           const double* __restrict__ sb,
                                                           const double* __restrict__ bprev,
                                                           const int* __restrict__ alive, double* __restrict__ normp,
                                                           const int use_prev) {
  using V2 = typename V2Of<T>::t;
  const int nthr = blockDim.x;  // multiple of R
  const int r = threadIdx.x % R;
  const bool live = alive[r] != 0;
  const double zr = z[r].x, zi = z[r].y, qa = sa[r], qb = bprev[r] * sb[r];
  double nrm = 0.0;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * nthr + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * nthr) {
    const V2 c = C[f], a = A[f];
    double wr = static_cast<double>(c.x), wi = static_cast<double>(c.y);
    const double ar = qa * static_cast<double>(a.x), ai = qa * static_cast<double>(a.y);
    wr -= zr * ar - zi * ai;  // w -= z * q
    wi -= zr * ai + zi * ar;
    if (use_prev) {  // w -= beta_prev * q_prev (skipped at k = 0, src/lanczos.cpp:186-190)
      const V2 b = B[f];
      wr -= qb * static_cast<double>(b.x);
      wi -= qb * static_cast<double>(b.y);
    }
    V2 out;
    out.x = static_cast<T>(live ? wr : 0.0);
    out.y = static_cast<T>(live ? wi : 0.0);
    C[f] = out;
    const double sr = static_cast<double>(out.x), si = static_cast<double>(out.y);
    nrm += sr * sr + si * si;
  }
  __shared__ double red[kUpdBlock];
  red[threadIdx.x] = nrm;
  __syncthreads();
  if (threadIdx.x < R) {
    double s = 0.0;
    for (int t = threadIdx.x; t < nthr; t += R) s += red[t];
    normp[static_cast<int64_t>(blockIdx.x) * R + threadIdx.x] = s;
  }
}

// The same update on 16-byte units with two units in flight per thread: a unit is one complex128
// value, or two complex64 values of vectors (2p, 2p + 1) (R even).  The block size is a multiple
// of P = R / EPU units and so is the grid stride, so each thread always sees the same vectors; it
// keeps one norm partial per vector of its unit.  Per value the arithmetic is k_update_flat's.
template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_update_units(int64_t nu, int R, typename V2Of<T>::t* __restrict__ C,
                                                            const typename V2Of<T>::t* __restrict__ A,
                                                            const typename V2Of<T>::t* __restrict__ B,
                                                            const double2* __restrict__ z, const double* __restrict__ sa,
                                                            const double* __restrict__ sb,
                                                            const double* __restrict__ bprev,
                                                            const int* __restrict__ alive, double* __restrict__ normp,
                                                            const int use_prev) {
  using V2 = typename V2Of<T>::t;
  constexpr int EPU = 16 / static_cast<int>(sizeof(V2));  // values per unit
  struct alignas(16) Unit {
    V2 v[EPU];
  };
  const int P = R / EPU;
  const int nthr = blockDim.x;  // multiple of P
  const int p = threadIdx.x % P;
  bool live[EPU];
  double zr[EPU], zi[EPU], qa[EPU], qb[EPU], nrm[EPU];
#pragma unroll
  for (int e = 0; e < EPU; ++e) {
    const int r = p * EPU + e;
    live[e] = alive[r] != 0;
    zr[e] = z[r].x;
    zi[e] = z[r].y;
    qa[e] = sa[r];
    qb[e] = bprev[r] * sb[r];
    nrm[e] = 0.0;
  }
  Unit* __restrict__ Cu = reinterpret_cast<Unit*>(C);
  const Unit* __restrict__ Au = reinterpret_cast<const Unit*>(A);
  const Unit* __restrict__ Bu = reinterpret_cast<const Unit*>(B);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * nthr;
  auto one = [&](const Unit& c, const Unit& a, const Unit& b, Unit& out) {
#pragma unroll
    for (int e = 0; e < EPU; ++e) {
      double wr = static_cast<double>(c.v[e].x), wi = static_cast<double>(c.v[e].y);
      const double ar = qa[e] * static_cast<double>(a.v[e].x), ai = qa[e] * static_cast<double>(a.v[e].y);
      wr -= zr[e] * ar - zi[e] * ai;
      wi -= zr[e] * ai + zi[e] * ar;
      if (use_prev) {
        wr -= qb[e] * static_cast<double>(b.v[e].x);
        wi -= qb[e] * static_cast<double>(b.v[e].y);
      }
      out.v[e].x = static_cast<T>(live[e] ? wr : 0.0);
      out.v[e].y = static_cast<T>(live[e] ? wi : 0.0);
      const double sr = static_cast<double>(out.v[e].x), si = static_cast<double>(out.v[e].y);
      nrm[e] += sr * sr + si * si;
    }
  };
  int64_t f = static_cast<int64_t>(blockIdx.x) * nthr + threadIdx.x;
  for (; f + stride < nu; f += 2 * stride) {  // two units in flight
    const Unit c0 = Cu[f], c1 = Cu[f + stride], a0 = Au[f], a1 = Au[f + stride];
    Unit b0{}, b1{};
    if (use_prev) {
      b0 = Bu[f];
      b1 = Bu[f + stride];
    }
    Unit o0, o1;
    one(c0, a0, b0, o0);
    one(c1, a1, b1, o1);
    Cu[f] = o0;
    Cu[f + stride] = o1;
  }
  if (f < nu) {
    const Unit c0 = Cu[f], a0 = Au[f];
    Unit b0{};
    if (use_prev) b0 = Bu[f];
    Unit o0;
    one(c0, a0, b0, o0);
    Cu[f] = o0;
  }
  __shared__ double red[EPU][kUpdBlock];
#pragma unroll
  for (int e = 0; e < EPU; ++e) red[e][threadIdx.x] = nrm[e];
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < R) {
    const int q = threadIdx.x / EPU, e = threadIdx.x % EPU;
    double s = 0.0;
    for (int t = q; t < nthr; t += P) s += red[e][t];
    normp[static_cast<int64_t>(blockIdx.x) * R + threadIdx.x] = s;
  }
}

// |q_r|^2 partials of the start block, same flat layout.
template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_norm_flat(int64_t n, int R, const typename V2Of<T>::t* __restrict__ q,
                                                         double* __restrict__ normp) {
  const int nthr = blockDim.x;
  double nrm = 0.0;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * nthr + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * nthr) {
    const auto v = q[f];
    nrm += static_cast<double>(v.x) * static_cast<double>(v.x) + static_cast<double>(v.y) * static_cast<double>(v.y);
  }
  __shared__ double red[kUpdBlock];
  red[threadIdx.x] = nrm;
  __syncthreads();
  if (threadIdx.x < R) {
    double s = 0.0;
    for (int t = threadIdx.x; t < nthr; t += R) s += red[t];
    normp[static_cast<int64_t>(blockIdx.x) * R + threadIdx.x] = s;
  }
}

// Block r: nw2 = sum_b normp[b][r]; stop or record beta and rotate the scales.
__global__ __launch_bounds__(kRedBlock) void k_reduce_beta(const double* __restrict__ normp, int nblk, int R,
                                                           int* __restrict__ alive, int* __restrict__ used,
                                                           double* __restrict__ beta, double* __restrict__ sa,
                                                           double* __restrict__ sb, double* __restrict__ bprev, int k,
                                                           double* __restrict__ sout) {
  const int r = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) a += normp[static_cast<int64_t>(i) * R + r];
  __shared__ double sh[kRedBlock / 64];
  a = wsum(a);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double nw2 = 0.0;
    for (int w = 0; w < kRedBlock / 64; ++w) nw2 += sh[w];
    if (alive[r]) {
      used[r] = k + 1;
      if (nw2 < 1e-24) {
        alive[r] = 0;
        sa[r] = 0.0;
        sb[r] = 0.0;
        bprev[r] = 0.0;
      } else {
        const double b = sqrt(nw2);
        beta[static_cast<int64_t>(k) * R + r] = b;
        sb[r] = sa[r];
        sa[r] = 1.0 / b;
        bprev[r] = b;
      }
    }
    if (sout) sout[r] = alive[r] ? sa[r] : 0.0;  // scale of the new Krylov vector (reorthogonalising runs)
  }
}

// sa_r = 1/|q_r| (or, when |q_r| < 1e-18, q_r[0] = 1 and sa_r = 1: src/lanczos.cpp:170-175).
__global__ void k_start_norm(const double* __restrict__ normp, int nblk, int R, double* sa) {
  const int r = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) a += normp[static_cast<int64_t>(i) * R + r];
  __shared__ double sh[kRedBlock / 64];
  a = wsum(a);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kRedBlock / 64; ++w) s += sh[w];
    const double n0 = sqrt(s);
    sa[r] = n0 < 1e-18 ? -1.0 : 1.0 / n0;  // -1 flags the unit-vector fallback
  }
}

template <typename T>
__global__ void k_start_fallback(typename V2Of<T>::t* q, int R, double* sa) {
  const int r = threadIdx.x;
  if (r < R && sa[r] < 0.0) {
    q[r].x = static_cast<T>(1.0);
    q[r].y = static_cast<T>(0.0);
    sa[r] = 1.0;
  }
}

// One sample's start, drawn contiguous (complex128 on the host), into slot r of the interleaved
// start block q[i * R + r] (converted to T).
template <typename T>
__global__ void k_sample_to_slot(int64_t n, int64_t first, int R, int r, const double2* __restrict__ src,
                                 typename V2Of<T>::t* __restrict__ q) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double2 z = src[i];
    q[(first + i) * R + r].x = static_cast<T>(z.x);
    q[(first + i) * R + r].y = static_cast<T>(z.y);
  }
}

// Random-phase start vectors e^{i 2 pi u}, u from a counter-based hash of (seed, sample, index).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ void k_phase_start(int64_t dim, int R, uint64_t seed, const uint32_t* sample_ids, typename V2Of<T>::t* q) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < dim;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    for (int r = 0; r < R; ++r) {
      const uint64_t h = mix64(mix64(seed * 0x9e3779b97f4a7c15ull + sample_ids[r]) + static_cast<uint64_t>(e));
      const double th = 6.283185307179586 * static_cast<double>(h >> 11) * 0x1.0p-53;
      double s, c;
      sincos(th, &s, &c);
      q[e * R + r].x = static_cast<T>(c);
      q[e * R + r].y = static_cast<T>(s);
    }
}

// The same start on a momentum stack: slot (j, r) of orbit g holds the value k_phase_start gives
// block j's row i = map_j[g] for sample id sample_ids[j * rk + r], and 0 where g is not a row of
// block j.
template <typename T>
__global__ void k_kstack_phase_start(int64_t nrep, int nk, int rk, KStackMaps maps, uint64_t seed,
                                     const uint32_t* sample_ids, typename V2Of<T>::t* q) {
  const int W = nk * rk;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < nrep;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x)
    for (int j = 0; j < nk; ++j) {
      const int32_t i = maps.map[j][g];
      for (int r = 0; r < rk; ++r) {
        T re = static_cast<T>(0), im = static_cast<T>(0);
        if (i >= 0) {
          const uint64_t h =
              mix64(mix64(seed * 0x9e3779b97f4a7c15ull + sample_ids[j * rk + r]) + static_cast<uint64_t>(i));
          const double th = 6.283185307179586 * static_cast<double>(h >> 11) * 0x1.0p-53;
          double s, c;
          sincos(th, &s, &c);
          re = static_cast<T>(c);
          im = static_cast<T>(s);
        }
        q[g * W + j * rk + r].x = re;
        q[g * W + j * rk + r].y = im;
      }
    }
}

void kstack_phase_start(const ftlm_op* stack, int rk, uint64_t seed, const uint32_t* sample_ids, int dtype, void* q,
                        hipStream_t st) {
  const int64_t n = stack->dim;
  if (n <= 0) return;
  const int nk = kstack_blocks(stack);
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(8192, (n + 255) / 256));
  if (dtype == FTLM_C64)
    hipLaunchKernelGGL(k_kstack_phase_start<float>, dim3(g), dim3(256), 0, st, n, nk, rk, kstack_maps(stack), seed,
                       sample_ids, static_cast<float2*>(q));
  else
    hipLaunchKernelGGL(k_kstack_phase_start<double>, dim3(g), dim3(256), 0, st, n, nk, rk, kstack_maps(stack), seed,
                       sample_ids, static_cast<double2*>(q));
  check_launch("k_kstack_phase_start");
}

// ---------------------------------------------------------------- reorthogonalisation
// With reorthogonalisation every Krylov block stays resident: slot j holds w_j (unnormalised, as
// the three-buffer recurrence leaves it) and scl[j][r] = 1/beta_{j-1} makes q_j = scl_j * slot_j.
// A pass computes h_j = <q_j|w> for j <= k (FP64), then w -= sum_j h_j q_j; a second pass, run
// when the first one cancelled much of w, makes classical Gram-Schmidt as accurate as the modified
// form (Giraud, Langou, Rozloznik 2005).
// flags[r] selects the vectors reorthogonalised this step, flags[R + r] forces the next step.
constexpr int kRoJ = 8;  // Krylov slots whose loads are in flight together

__device__ __forceinline__ bool ro_any(const int* __restrict__ flags, int R) {
  int any = 0;
  for (int r = 0; r < R; ++r) any |= flags[r];
  return any != 0;
}

// Per-workgroup partials of <slot_j|w> * scl_j for j < nq: part[blk][j][r] (row stride nq_cap).
template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_ro_dots(int64_t n, int R, const typename V2Of<T>::t* __restrict__ C,
                                                       const typename V2Of<T>::t* __restrict__ slots, int64_t stride,
                                                       int nq, int nq_cap, const double* __restrict__ scl,
                                                       const int* __restrict__ flags, double2* __restrict__ part) {
  using V2 = typename V2Of<T>::t;
  if (!ro_any(flags, R)) return;
  const int nthr = blockDim.x, r = threadIdx.x % R;
  __shared__ double red[2][kUpdBlock];
  for (int j0 = 0; j0 < nq; j0 += kRoJ) {
    double ar[kRoJ], ai[kRoJ];
#pragma unroll
    for (int jj = 0; jj < kRoJ; ++jj) ar[jj] = ai[jj] = 0.0;
    for (int64_t f = static_cast<int64_t>(blockIdx.x) * nthr + threadIdx.x; f < n;
         f += static_cast<int64_t>(gridDim.x) * nthr) {
      const V2 c = C[f];
      const double cr = static_cast<double>(c.x), ci = static_cast<double>(c.y);
      V2 q[kRoJ];
#pragma unroll
      for (int jj = 0; jj < kRoJ; ++jj)
        if (j0 + jj < nq) q[jj] = slots[static_cast<int64_t>(j0 + jj) * stride + f];
#pragma unroll
      for (int jj = 0; jj < kRoJ; ++jj)
        if (j0 + jj < nq) {
          const double qr = static_cast<double>(q[jj].x), qi = static_cast<double>(q[jj].y);
          ar[jj] += qr * cr + qi * ci;  // conj(q) * w
          ai[jj] += qr * ci - qi * cr;
        }
    }
    for (int jj = 0; jj < kRoJ && j0 + jj < nq; ++jj) {
      red[0][threadIdx.x] = ar[jj];
      red[1][threadIdx.x] = ai[jj];
      __syncthreads();
      if (threadIdx.x < R) {
        double x = 0.0, y = 0.0;
        for (int t = threadIdx.x; t < nthr; t += R) {
          x += red[0][t];
          y += red[1][t];
        }
        const double s = scl[static_cast<int64_t>(j0 + jj) * R + threadIdx.x];
        part[(static_cast<int64_t>(blockIdx.x) * nq_cap + j0 + jj) * R + threadIdx.x] = make_double2(s * x, s * y);
      }
      __syncthreads();
    }
  }
}

// Block (j, r): coef[j][r] = scl_j * <q_j|w> for vectors being reorthogonalised, else 0.
__global__ __launch_bounds__(kRedBlock) void k_ro_reduce(const double2* __restrict__ part, int nblk, int nq_cap, int R,
                                                         const double* __restrict__ scl, const int* __restrict__ flags,
                                                         double2* __restrict__ coef) {
  if (!ro_any(flags, R)) return;
  const int j = blockIdx.x, r = blockIdx.y;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) {
    const double2 v = part[(static_cast<int64_t>(i) * nq_cap + j) * R + r];
    a += v.x;
    b += v.y;
  }
  __shared__ double sh[2][kRedBlock / 64];
  a = wsum(a);
  b = wsum(b);
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int w = 0; w < kRedBlock / 64; ++w) {
      x += sh[0][w];
      y += sh[1][w];
    }
    const double s = flags[r] ? scl[static_cast<int64_t>(j) * R + r] : 0.0;
    coef[static_cast<int64_t>(j) * R + r] = make_double2(s * x, s * y);
  }
}

// w -= sum_{j<nq} coef_j slot_j for flagged vectors, then |w|^2 partials of every vector in the
// update kernel's layout (same grid, block and element order, so unchanged vectors reproduce the
// update's partials bit for bit).
template <typename T>
__global__ __launch_bounds__(kUpdBlock) void k_ro_sub(int64_t n, int R, typename V2Of<T>::t* __restrict__ C,
                                                      const typename V2Of<T>::t* __restrict__ slots, int64_t stride,
                                                      int nq, const double2* __restrict__ coef,
                                                      const int* __restrict__ flags, double* __restrict__ normp) {
  using V2 = typename V2Of<T>::t;
  if (!ro_any(flags, R)) return;
  const int nthr = blockDim.x, r = threadIdx.x % R;
  const bool on = flags[r] != 0;
  double nrm = 0.0;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * nthr + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * nthr) {
    V2 c = C[f];
    if (on) {
      double wr = static_cast<double>(c.x), wi = static_cast<double>(c.y);
      for (int j0 = 0; j0 < nq; j0 += kRoJ) {
        V2 q[kRoJ];
#pragma unroll
        for (int jj = 0; jj < kRoJ; ++jj)
          if (j0 + jj < nq) q[jj] = slots[static_cast<int64_t>(j0 + jj) * stride + f];
#pragma unroll
        for (int jj = 0; jj < kRoJ; ++jj)
          if (j0 + jj < nq) {
            const double2 h = coef[static_cast<int64_t>(j0 + jj) * R + r];
            const double qr = static_cast<double>(q[jj].x), qi = static_cast<double>(q[jj].y);
            wr -= h.x * qr - h.y * qi;
            wi -= h.x * qi + h.y * qr;
          }
      }
      c.x = static_cast<T>(wr);
      c.y = static_cast<T>(wi);
      C[f] = c;
    }
    const double sr = static_cast<double>(c.x), si = static_cast<double>(c.y);
    nrm += sr * sr + si * si;
  }
  __shared__ double red[kUpdBlock];
  red[threadIdx.x] = nrm;
  __syncthreads();
  if (threadIdx.x < R) {
    double s = 0.0;
    for (int t = threadIdx.x; t < nthr; t += R) s += red[t];
    normp[static_cast<int64_t>(blockIdx.x) * R + threadIdx.x] = s;
  }
}

// Decides, per vector, whether step k reorthogonalises w_{k+1}.  FULL: every live vector.
// PARTIAL (H. D. Simon, Math. Comp. 42 (1984) 115-142): with beta_k = |w| before
// reorthogonalisation, the estimates omega_{k+1,j} of |<q_{k+1}|q_j>| follow
//   beta_k w_{k+1,j} = beta_j w_{k,j+1} + (alpha_j - alpha_k) w_{k,j} + beta_{j-1} w_{k,j-1}
//                      - beta_{k-1} w_{k-1,j} + sign * eps1 (beta_k + beta_j),
// with w_{k+1,k} = eps1 = eps sqrt(dim) and w_{k+1,k+1} = 1.  A step whose largest estimate
// exceeds sqrt(eps) is reorthogonalised, and so is the one after; the estimates of a
// reorthogonalised vector restart at eps1.  Rows live in om[3][R][ld], row (k mod 3) = omega_k.
__global__ __launch_bounds__(kRedBlock) void k_ro_omega(int k, int R, int mode, double eps, double eps1,
                                                        const double* __restrict__ alpha,
                                                        const double* __restrict__ beta,
                                                        const double* __restrict__ normp, int nblk,
                                                        const int* __restrict__ alive, double* __restrict__ om,
                                                        int ld, int* __restrict__ flags, double* __restrict__ nw0) {
  const int r = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) a += normp[static_cast<int64_t>(i) * R + r];
  __shared__ double sh[kRedBlock / 64];
  __shared__ int sflag;
  a = wsum(a);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  double nw2 = 0.0;
  for (int w = 0; w < kRedBlock / 64; ++w) nw2 += sh[w];
  const bool live = alive[r] != 0 && nw2 >= 1e-24;
  if (threadIdx.x == 0) nw0[r] = nw2;
  if (mode == FTLM_REORTH_FULL || !live) {
    if (threadIdx.x == 0) {
      flags[r] = live ? 1 : 0;
      flags[R + r] = 0;
    }
    return;
  }
  const double bk = sqrt(nw2);
  const double* wk = om + (static_cast<int64_t>(k % 3) * R + r) * ld;
  const double* wkm = om + (static_cast<int64_t>((k + 2) % 3) * R + r) * ld;
  double* wn = om + (static_cast<int64_t>((k + 1) % 3) * R + r) * ld;
  const double ak = alpha[static_cast<int64_t>(k) * R + r];
  const double bkm = k > 0 ? beta[static_cast<int64_t>(k - 1) * R + r] : 0.0;
  double mx = 0.0;
  for (int j = threadIdx.x; j < k; j += kRedBlock) {
    const double bj = beta[static_cast<int64_t>(j) * R + r];
    double v = bj * wk[j + 1] + (alpha[static_cast<int64_t>(j) * R + r] - ak) * wk[j] - bkm * wkm[j];
    if (j > 0) v += beta[static_cast<int64_t>(j - 1) * R + r] * wk[j - 1];
    v = (v + copysign(eps1 * (bk + bj), v)) / bk;
    wn[j] = v;
    mx = fmax(mx, fabs(v));
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kRedBlock / 64; ++w) m = fmax(m, sh[w]);
    const int trig = m > sqrt(eps) ? 1 : 0;
    const int on = trig | flags[R + r];
    flags[r] = on;
    flags[R + r] = trig;
    sflag = on;
    wn[k] = eps1;
    wn[k + 1] = 1.0;
  }
  __syncthreads();
  if (sflag)
    for (int j = threadIdx.x; j < k; j += kRedBlock) wn[j] = eps1;
}

// After the first Gram-Schmidt pass: keep a vector for the second pass only when the pass cut
// its norm below 1/sqrt(2) of what it was (the criterion of Daniel, Gragg, Kaufman and Stewart,
// Math. Comp. 30 (1976) 772-795); otherwise one pass already left it orthogonal to rounding.
__global__ __launch_bounds__(kRedBlock) void k_ro_second(const double* __restrict__ normp, int nblk, int R,
                                                         const double* __restrict__ nw0, int* __restrict__ flags) {
  const int r = blockIdx.x;
  if (!flags[r]) return;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedBlock) a += normp[static_cast<int64_t>(i) * R + r];
  __shared__ double sh[kRedBlock / 64];
  a = wsum(a);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double nw1 = 0.0;
    for (int w = 0; w < kRedBlock / 64; ++w) nw1 += sh[w];
    flags[r] = nw1 < 0.5 * nw0[r] ? 1 : 0;
  }
}

// omega_0 = (1), flags cleared: the start of a reorthogonalising recurrence.
__global__ void k_ro_init(int R, double* __restrict__ om, int64_t om_n, int ld, int* __restrict__ flags) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < om_n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    om[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x < 2 * R) flags[threadIdx.x] = 0;
}

__global__ void k_ro_row0(int R, double* __restrict__ om, int ld) {
  if (threadIdx.x < R) om[static_cast<int64_t>(threadIdx.x) * ld] = 1.0;
}

// ---------------------------------------------------------------- host engine

struct Krylov {
  const ftlm_op* op = nullptr;
  int R = 1, M = 1, dtype = FTLM_C128;
  int64_t dim = 0;
  size_t vbytes = 0;  // bytes of one block of R vectors
  DevBuf<unsigned char> buf[3];
  DevBuf<double> sa, sb, bprev, alpha, beta;
  DevBuf<double2> z, dotp;
  DevBuf<double> normp;
  DevBuf<int> alive, used;
  int nblk_hv = 0, nblk_upd = 0;
  bool force_real = false;
  // reorthogonalisation (FTLM_REORTH_*): every Krylov block in `basis` (slot j = w_j), per-slot
  // scales scl[M+1][R], overlap partials/coefficients, omega rows and the step flags
  int reorth = FTLM_REORTH_NONE;
  int nblk_dot = 0, om_ld = 0;
  DevBuf<unsigned char> basis;
  DevBuf<double> scl, om, rnw0;
  DevBuf<double2> rpart, rcoef;
  DevBuf<int> rflags;

  void init(const ftlm_op* o, int R_, int M_, int dt, int reorth_ = FTLM_REORTH_NONE) {
    op = o;
    R = R_;
    M = M_;
    dtype = dt;
    dim = o->dim;
    reorth = reorth_;
    vbytes = static_cast<size_t>(dim) * R * (dt == FTLM_C64 ? sizeof(float2) : sizeof(double2));
    if (reorth == FTLM_REORTH_NONE) {
      for (auto& b : buf) b.alloc(vbytes);
    } else {
      basis.alloc(static_cast<size_t>(M + 1) * vbytes);
      scl.alloc(static_cast<size_t>(M + 1) * R);
      rcoef.alloc(static_cast<size_t>(M + 1) * R);
      rflags.alloc(static_cast<size_t>(2) * R);
      rnw0.alloc(static_cast<size_t>(R));
      om_ld = M + 2;
      if (reorth == FTLM_REORTH_PARTIAL) om.alloc(static_cast<size_t>(3) * R * om_ld);
    }
    sa.alloc(R);
    sb.alloc(R);
    bprev.alloc(R);
    alpha.alloc(static_cast<size_t>(M) * R);
    beta.alloc(static_cast<size_t>(M) * R);
    z.alloc(R);
    alive.alloc(R);
    used.alloc(R);
    nblk_hv = hv_partial_blocks(o);
    nblk_upd = static_cast<int>(std::min<int64_t>(kUpdMaxGrid, (dim * R + upd_threads() - 1) / upd_threads()));
    dotp.alloc(static_cast<size_t>(nblk_hv) * R);
    normp.alloc(static_cast<size_t>(std::max(nblk_hv, nblk_upd)) * R);
    if (reorth != FTLM_REORTH_NONE) {
      nblk_dot = std::min(nblk_upd, 1024);
      rpart.alloc(static_cast<size_t>(nblk_dot) * (M + 1) * R);
    }
  }

  // Krylov block j of a reorthogonalising run; the start block otherwise.
  void* slot(int j) const { return reorth == FTLM_REORTH_NONE ? static_cast<void*>(buf[0].p) : basis.p + j * vbytes; }

  // block size: the largest multiple of R not above kUpdBlock (each thread keeps one vector r)
  int upd_threads() const { return (kUpdBlock / R) * R; }

  void launch_update(void* C, const void* A, const void* B, int use_prev, hipStream_t st) {
    const int64_t n = dim * R;
    if (dtype == FTLM_C128 || R % 2 == 0) {  // 16-byte units (k_update_units)
      const int epu = dtype == FTLM_C64 ? 2 : 1, P = R / epu;
      const int thr = (kUpdBlock / P) * P;
      const int64_t nu = n / epu;
      const dim3 g(nblk_upd), b(thr);  // nblk_upd partial blocks, as the readers of normp expect
      if (dtype == FTLM_C64)
        hipLaunchKernelGGL(k_update_units<float>, g, b, 0, st, nu, R, static_cast<float2*>(C),
                           static_cast<const float2*>(A), static_cast<const float2*>(B), z.p, sa.p, sb.p, bprev.p,
                           alive.p, normp.p, use_prev);
      else
        hipLaunchKernelGGL(k_update_units<double>, g, b, 0, st, nu, R, static_cast<double2*>(C),
                           static_cast<const double2*>(A), static_cast<const double2*>(B), z.p, sa.p, sb.p, bprev.p,
                           alive.p, normp.p, use_prev);
      check_launch("k_update_units");
      return;
    }
    const dim3 g(nblk_upd), b(upd_threads());
    if (dtype == FTLM_C64)
      hipLaunchKernelGGL(k_update_flat<float>, g, b, 0, st, n, R, static_cast<float2*>(C),
                         static_cast<const float2*>(A), static_cast<const float2*>(B), z.p, sa.p, sb.p, bprev.p,
                         alive.p, normp.p, use_prev);
    else
      hipLaunchKernelGGL(k_update_flat<double>, g, b, 0, st, n, R, static_cast<double2*>(C),
                         static_cast<const double2*>(A), static_cast<const double2*>(B), z.p, sa.p, sb.p, bprev.p,
                         alive.p, normp.p, use_prev);
    check_launch("k_update_flat");
  }

  // Prepare the start block (buf[0], or slot 0): normalise lazily and reset the recurrence state.
  void start(hipStream_t st) {
    std::vector<int> one(static_cast<size_t>(R), 1), zero(static_cast<size_t>(R), 0);
    std::vector<double> dz(static_cast<size_t>(R), 0.0);
    alive.upload(one, st);
    used.upload(zero, st);
    sb.upload(dz, st);
    bprev.upload(dz, st);
    FTLM_HIP_CHECK(hipMemsetAsync(alpha.p, 0, alpha.n * sizeof(double), st));
    FTLM_HIP_CHECK(hipMemsetAsync(beta.p, 0, beta.n * sizeof(double), st));
    const dim3 g(nblk_upd), b(upd_threads());
    if (dtype == FTLM_C64)
      hipLaunchKernelGGL(k_norm_flat<float>, g, b, 0, st, dim * R, R, static_cast<const float2*>((void*)slot(0)),
                         normp.p);
    else
      hipLaunchKernelGGL(k_norm_flat<double>, g, b, 0, st, dim * R, R, static_cast<const double2*>((void*)slot(0)),
                         normp.p);
    hipLaunchKernelGGL(k_start_norm, dim3(R), dim3(kRedBlock), 0, st, normp.p, nblk_upd, R, sa.p);
    if (dtype == FTLM_C64)
      hipLaunchKernelGGL(k_start_fallback<float>, dim3(1), dim3(64), 0, st, static_cast<float2*>((void*)slot(0)), R, sa.p);
    else
      hipLaunchKernelGGL(k_start_fallback<double>, dim3(1), dim3(64), 0, st, static_cast<double2*>((void*)slot(0)), R,
                         sa.p);
    check_launch("start");
    start_reorth(st);
  }

  // Host-computed start norms (bit-identical to the reference's sequential sum).
  void start_host_norm(const std::vector<double>& inv_norm, hipStream_t st) {
    std::vector<int> one(static_cast<size_t>(R), 1), zero(static_cast<size_t>(R), 0);
    std::vector<double> dz(static_cast<size_t>(R), 0.0);
    alive.upload(one, st);
    used.upload(zero, st);
    sb.upload(dz, st);
    bprev.upload(dz, st);
    sa.upload(inv_norm, st);
    FTLM_HIP_CHECK(hipMemsetAsync(alpha.p, 0, alpha.n * sizeof(double), st));
    FTLM_HIP_CHECK(hipMemsetAsync(beta.p, 0, beta.n * sizeof(double), st));
    start_reorth(st);
  }

  // One Lanczos step k with (A, B, C) = buffers (a, b, c).
  void step(int k, int a, int b, int c, hipStream_t st) {
    const int nb = hv_apply(op, buf[a].p, buf[c].p, R, dtype, sa.p, dotp.p, st, force_real);
    hipLaunchKernelGGL(k_reduce_alpha, dim3(R), dim3(kRedBlock), 0, st, dotp.p, nb, R, alive.p, z.p, alpha.p, k);
    launch_update(buf[c].p, buf[a].p, buf[b].p, k > 0 ? 1 : 0, st);
    hipLaunchKernelGGL(k_reduce_beta, dim3(R), dim3(kRedBlock), 0, st, normp.p, nblk_upd, R, alive.p, used.p, beta.p,
                       sa.p, sb.p, bprev.p, k, static_cast<double*>(nullptr));
    check_launch("lanczos step");
  }

  // Clears the reorthogonalisation state after the start block (slot 0, scale sa) is in place.
  void start_reorth(hipStream_t st) {
    if (reorth == FTLM_REORTH_NONE) return;
    FTLM_HIP_CHECK(hipMemcpyAsync(scl.p, sa.p, sizeof(double) * R, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_ro_init, dim3(om.n ? static_cast<unsigned>(std::min<size_t>(1024, (om.n + 255) / 256)) : 1),
                       dim3(256), 0, st, R, om.p, static_cast<int64_t>(om.n), om_ld, rflags.p);
    if (om.n) hipLaunchKernelGGL(k_ro_row0, dim3(1), dim3(64), 0, st, R, om.p, om_ld);
    check_launch("reorthogonalisation start");
  }

  // One step of a reorthogonalising run: A = slot k, B = slot k-1, C = slot k+1.
  void step_reorth(int k, hipStream_t st) {
    void* A = slot(k);
    void* B = slot(k > 0 ? k - 1 : k);
    void* C = slot(k + 1);
    const int nb = hv_apply(op, A, C, R, dtype, sa.p, dotp.p, st, force_real);
    hipLaunchKernelGGL(k_reduce_alpha, dim3(R), dim3(kRedBlock), 0, st, dotp.p, nb, R, alive.p, z.p, alpha.p, k);
    launch_update(C, A, B, k > 0 ? 1 : 0, st);
    const double eps = dtype == FTLM_C64 ? 0x1.0p-24 : 0x1.0p-53;
    const double eps1 = eps * std::sqrt(static_cast<double>(dim));
    hipLaunchKernelGGL(k_ro_omega, dim3(R), dim3(kRedBlock), 0, st, k, R, reorth, eps, eps1, alpha.p, beta.p, normp.p,
                       nblk_upd, alive.p, om.p, om_ld, rflags.p, rnw0.p);
    const int64_t n = dim * R;
    const int nq = k + 1;
    for (int pass = 0; pass < 2; ++pass) {
      if (dtype == FTLM_C64)
        hipLaunchKernelGGL(k_ro_dots<float>, dim3(nblk_dot), dim3(upd_threads()), 0, st, n, R,
                           static_cast<const float2*>(C), static_cast<const float2*>((void*)basis.p), n, nq, M + 1,
                           scl.p, rflags.p, rpart.p);
      else
        hipLaunchKernelGGL(k_ro_dots<double>, dim3(nblk_dot), dim3(upd_threads()), 0, st, n, R,
                           static_cast<const double2*>(C), static_cast<const double2*>((void*)basis.p), n, nq, M + 1,
                           scl.p, rflags.p, rpart.p);
      hipLaunchKernelGGL(k_ro_reduce, dim3(nq, R), dim3(kRedBlock), 0, st, rpart.p, nblk_dot, M + 1, R, scl.p,
                         rflags.p, rcoef.p);
      if (dtype == FTLM_C64)
        hipLaunchKernelGGL(k_ro_sub<float>, dim3(nblk_upd), dim3(upd_threads()), 0, st, n, R, static_cast<float2*>(C),
                           static_cast<const float2*>((void*)basis.p), n, nq, rcoef.p, rflags.p, normp.p);
      else
        hipLaunchKernelGGL(k_ro_sub<double>, dim3(nblk_upd), dim3(upd_threads()), 0, st, n, R,
                           static_cast<double2*>(C), static_cast<const double2*>((void*)basis.p), n, nq, rcoef.p,
                           rflags.p, normp.p);
      if (pass == 0)
        hipLaunchKernelGGL(k_ro_second, dim3(R), dim3(kRedBlock), 0, st, normp.p, nblk_upd, R, rnw0.p, rflags.p);
    }
    hipLaunchKernelGGL(k_reduce_beta, dim3(R), dim3(kRedBlock), 0, st, normp.p, nblk_upd, R, alive.p, used.p, beta.p,
                       sa.p, sb.p, bprev.p, k, scl.p + static_cast<int64_t>(k + 1) * R);
    check_launch("reorthogonalising lanczos step");
  }

  // Steps [k0, k1); buffer roles rotate (A,B,C) -> (C,A,B) every step, starting from A = buf[0].
  void run(int k0, int k1, hipStream_t st) {
    for (int k = k0; k < k1; ++k) {
      if (reorth != FTLM_REORTH_NONE) {
        step_reorth(k, st);
        continue;
      }
      const int a = (3 - k % 3) % 3;  // k=0 -> 0, k=1 -> 2, k=2 -> 1
      const int b = (a + 1) % 3, c = (a + 2) % 3;
      step(k, a, b, c, st);
    }
  }

  // The same steps captured as one hipGraph on a private stream (ordered after `st` and before
  // anything later on `st`) and launched once.  Every step is kernel launches only: the H*V
  // tile counters reset themselves and partial sums are cleared by a kernel, never a memset node.
  void run_graph(int k0, int k1, hipStream_t st) {
    FTLM_REQUIRE(op->kind != kCallback && op->kind != kGram, FTLM_EUNSUPPORTED,
                 "host-callback operators and momentum Gram blocks cannot be graph-captured");
    if (k1 <= k0) return;
    hipStream_t cs = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    auto cleanup = [&] {
      if (ex) (void)hipGraphExecDestroy(ex);
      if (g) (void)hipGraphDestroy(g);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (cs) (void)hipStreamDestroy(cs);
    };
    try {
      FTLM_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      FTLM_HIP_CHECK(hipEventCreateWithFlags(&e0, hipEventDisableTiming));
      FTLM_HIP_CHECK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
      FTLM_HIP_CHECK(hipEventRecord(e0, st));
      FTLM_HIP_CHECK(hipStreamWaitEvent(cs, e0, 0));
      FTLM_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      try {
        run(k0, k1, cs);
      } catch (...) {
        (void)hipStreamEndCapture(cs, &g);
        throw;
      }
      FTLM_HIP_CHECK(hipStreamEndCapture(cs, &g));
      FTLM_HIP_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      FTLM_HIP_CHECK(hipGraphLaunch(ex, cs));
      FTLM_HIP_CHECK(hipEventRecord(e1, cs));
      FTLM_HIP_CHECK(hipStreamWaitEvent(st, e1, 0));
      FTLM_HIP_CHECK(hipStreamSynchronize(cs));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  }
};

// Reference start vector: mt19937(seed) + normal(0,1), (re, im) per index.
void gauss_mt(int64_t dim, unsigned seed, std::vector<std::complex<double>>* q) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> g(0.0, 1.0);
  q->resize(static_cast<size_t>(dim));
  for (int64_t i = 0; i < dim; ++i) {
    const double re = g(rng);
    const double im = g(rng);
    (*q)[static_cast<size_t>(i)] = std::complex<double>(re, im);
  }
}

// ---- the reference starts of many samples, drawn in parallel
// Each sample is one sequential mt19937 / normal_distribution stream (gauss_mt), so the samples are
// drawn in parallel, one thread per sample (up to 16), in blocks of kGaussBlock indices.  A
// sample's generator and distribution carry over between blocks, so it sees exactly the serial
// sequence.
constexpr int64_t kGaussBlock = int64_t(1) << 18;

static int gauss_threads(int R) {
  return std::max(1, std::min(R, std::min(16, static_cast<int>(std::thread::hardware_concurrency()))));
}

// body(0 .. n-1), one call per thread; the calling thread runs body(0).
static void on_threads(int n, const std::function<void(int)>& body) {
  std::vector<std::thread> th;
  for (int t = 1; t < n; ++t) th.emplace_back(body, t);
  body(0);
  for (auto& x : th) x.join();
}

struct GaussStreams {
  std::vector<std::mt19937> rng;
  std::vector<std::normal_distribution<double>> g;
  GaussStreams(const std::vector<unsigned>& seeds, int R)
      : g(static_cast<size_t>(R), std::normal_distribution<double>(0.0, 1.0)) {
    for (int r = 0; r < R; ++r) rng.emplace_back(seeds[static_cast<size_t>(r)]);
  }
  // The next nb values of sample r.  The drawing thread works on its own copies: the objects sit
  // side by side in the vectors, and drawing in place would share cache lines between threads.
  void draw(int r, int64_t nb, std::complex<double>* out) {
    std::mt19937 e = rng[static_cast<size_t>(r)];
    std::normal_distribution<double> d = g[static_cast<size_t>(r)];
    for (int64_t i = 0; i < nb; ++i) {
      const double re = d(e);
      const double im = d(e);
      out[i] = std::complex<double>(re, im);
    }
    rng[static_cast<size_t>(r)] = e;
    g[static_cast<size_t>(r)] = d;
  }
};

// Large starts of few samples: each sample's one stream is split over the host threads with
// MT19937 jump-ahead (gauss_mt.hpp), exactly; with many samples, one thread per sample is as fast.
constexpr int64_t kSplitGaussMinDim = int64_t(1) << 21;

static int split_gauss_threads() {
  return std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
}

static bool split_gauss(int64_t dim, int R) { return dim >= kSplitGaussMinDim && 2 * R <= split_gauss_threads(); }

// The reference starts of R samples, unnormalised and interleaved: h[i * R + r] = sample r at i.
// Each block is drawn per sample into private buffers, then the threads interleave disjoint index
// ranges.
void gauss_interleaved(int64_t dim, int R, const std::vector<unsigned>& seeds, std::vector<double2>* h) {
  h->resize(static_cast<size_t>(dim) * R);
  if (dim <= 0) return;
  if (split_gauss(dim, R)) {
    for (int r = 0; r < R; ++r)
      ftlm_gauss::draw_sample(seeds[static_cast<size_t>(r)], dim, split_gauss_threads(),
                              [&](int64_t first, int64_t n, const std::complex<double>* v) {
                                for (int64_t i = 0; i < n; ++i)
                                  (*h)[static_cast<size_t>(first + i) * R + r] = make_double2(v[i].real(), v[i].imag());
                              });
    return;
  }
  const int nthr = gauss_threads(R);
  GaussStreams gs(seeds, R);
  const int64_t blk = std::min(kGaussBlock, dim);
  std::vector<std::vector<std::complex<double>>> part(static_cast<size_t>(R),
                                                      std::vector<std::complex<double>>(static_cast<size_t>(blk)));
  for (int64_t b0 = 0; b0 < dim; b0 += blk) {
    const int64_t nb = std::min(blk, dim - b0);
    on_threads(nthr, [&](int t) {
      for (int r = t; r < R; r += nthr) gs.draw(r, nb, part[static_cast<size_t>(r)].data());
    });
    on_threads(nthr, [&](int t) {
      const int64_t i0 = nb * t / nthr, i1 = nb * (t + 1) / nthr;
      for (int64_t i = i0; i < i1; ++i)
        for (int r = 0; r < R; ++r) {
          const std::complex<double> z = part[static_cast<size_t>(r)][static_cast<size_t>(i)];
          (*h)[static_cast<size_t>(b0 + i) * R + r] = make_double2(z.real(), z.imag());
        }
    });
  }
}

// The reference Gaussian starts of R samples, sample after sample, each drawn by all host threads
// (ftlm_gauss::draw_sample hands its entries over in index order): the calling thread sums the
// norm in index order (one serial sum, as the reference's norm2), packs the entries into one of
// two pinned chunks whose copy to a contiguous device block overlaps the next chunk, and a kernel
// moves the sample into its slot of the interleaved start block.
void finish_start(Krylov& K, const std::vector<double>& ss, hipStream_t st);

void fill_start_split_gauss(Krylov& K, const std::vector<unsigned>& seeds, hipStream_t st) {
  const int R = K.R;
  const int64_t dim = K.dim;
  constexpr int64_t kChunk = int64_t(1) << 21;  // entries per pinned chunk (32 MB)
  struct Pinned {
    double2* h[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Pinned() {
      for (int i = 0; i < 2; ++i) {
        if (ev[i]) {
          (void)hipEventSynchronize(ev[i]);
          (void)hipEventDestroy(ev[i]);
        }
        if (h[i]) (void)hipHostFree(h[i]);
      }
    }
  } pin;
  for (int i = 0; i < 2; ++i) {
    FTLM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&pin.h[i]), kChunk * sizeof(double2), hipHostMallocDefault));
    FTLM_HIP_CHECK(hipEventCreateWithFlags(&pin.ev[i], hipEventDisableTiming));
  }
  DevBuf<double2> dev(static_cast<size_t>(dim));
  std::vector<double> ss(static_cast<size_t>(R), 0.0);
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(8192, (dim + 255) / 256));
  for (int r = 0; r < R; ++r) {
    double s2 = 0.0;
    int64_t fill = 0, base = 0;
    int which = 0;
    bool used[2] = {false, false};
    auto flush = [&] {
      if (fill == 0) return;
      FTLM_HIP_CHECK(hipMemcpyAsync(dev.p + base, pin.h[which], fill * sizeof(double2), hipMemcpyHostToDevice, st));
      FTLM_HIP_CHECK(hipEventRecord(pin.ev[which], st));
      used[which] = true;
      base += fill;
      fill = 0;
      which ^= 1;
      if (used[which]) FTLM_HIP_CHECK(hipEventSynchronize(pin.ev[which]));  // its last copy is done
    };
    ftlm_gauss::draw_sample(seeds[static_cast<size_t>(r)], dim, split_gauss_threads(),
                            [&](int64_t, int64_t n, const std::complex<double>* v) {
                              for (int64_t i = 0; i < n; ++i) {
                                s2 += std::norm(v[i]);
                                pin.h[which][fill++] = make_double2(v[i].real(), v[i].imag());
                                if (fill == kChunk) flush();
                              }
                            });
    flush();
    ss[static_cast<size_t>(r)] = s2;
    if (K.dtype == FTLM_C64)
      hipLaunchKernelGGL(k_sample_to_slot<float>, dim3(g), dim3(256), 0, st, dim, int64_t{0}, R, r, dev.p,
                         static_cast<float2*>(K.slot(0)));
    else
      hipLaunchKernelGGL(k_sample_to_slot<double>, dim3(g), dim3(256), 0, st, dim, int64_t{0}, R, r, dev.p,
                         static_cast<double2*>(K.slot(0)));
    check_launch("k_sample_to_slot");
    FTLM_HIP_CHECK(hipStreamSynchronize(st));  // dev is reused by the next sample
  }
  finish_start(K, ss, st);
}

// Fill the start block with R start vectors.  seeds[r] per sample.
void fill_start(Krylov& K, int start, const std::vector<unsigned>& seeds, const double* q0, hipStream_t st) {
  const int R = K.R;
  const int64_t dim = K.dim;
  if (K.op->kind == kKStack) {  // slot (j, r) starts as block j's sample with seed seeds[j * rk + r]
    FTLM_REQUIRE(start == FTLM_START_PHASE && q0 == nullptr, FTLM_EUNSUPPORTED,
                 "a momentum stack takes random-phase starts only");
    const int nk = kstack_blocks(K.op);
    FTLM_REQUIRE(R % nk == 0, FTLM_EINVAL, "a momentum stack takes R = blocks * samples per block");
    const int rk = R / nk;
    std::vector<uint32_t> ids(static_cast<size_t>(R));
    for (int v = 0; v < R; ++v) ids[v] = seeds[v];
    DevBuf<uint32_t> dids(static_cast<size_t>(R));
    dids.upload(ids, st);
    kstack_phase_start(K.op, rk, 0x5eedull, dids.p, K.dtype, K.slot(0), st);
    K.start(st);
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  if (start == FTLM_START_PHASE && q0 == nullptr) {
    std::vector<uint32_t> ids(static_cast<size_t>(R));
    for (int r = 0; r < R; ++r) ids[r] = seeds[r];
    DevBuf<uint32_t> dids(static_cast<size_t>(R));
    dids.upload(ids, st);
    const unsigned g = static_cast<unsigned>(std::min<int64_t>(8192, (dim + 255) / 256));
    if (K.dtype == FTLM_C64)
      hipLaunchKernelGGL(k_phase_start<float>, dim3(g), dim3(256), 0, st, dim, R, 0x5eedull, dids.p,
                         static_cast<float2*>(K.slot(0)));
    else
      hipLaunchKernelGGL(k_phase_start<double>, dim3(g), dim3(256), 0, st, dim, R, 0x5eedull, dids.p,
                         static_cast<double2*>(K.slot(0)));
    check_launch("k_phase_start");
    K.start(st);
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
    return;
  }
  if (q0 == nullptr && split_gauss(dim, R)) {
    fill_start_split_gauss(K, seeds, st);
    return;
  }
  // host-generated (reference mt19937 Gaussian or caller-supplied), interleaved [dim][R], block by
  // block (GaussStreams): the threads interleave disjoint index ranges of each block into one of
  // two pinned staging blocks, whose copy to the device overlaps the next block's draws.  The norms
  // accumulate per sample in index order, as one serial sum.
  const size_t es = K.dtype == FTLM_C64 ? sizeof(float2) : sizeof(double2);
  const int nthr = gauss_threads(R);
  GaussStreams gs(q0 == nullptr ? seeds : std::vector<unsigned>(static_cast<size_t>(R), 0u), R);
  const int64_t blk = std::min(kGaussBlock, dim);
  std::vector<std::vector<std::complex<double>>> part(static_cast<size_t>(R),
                                                      std::vector<std::complex<double>>(static_cast<size_t>(blk)));
  std::vector<double> ss(static_cast<size_t>(R), 0.0);
  // A single block (dim <= kGaussBlock) goes through one pageable buffer and a synchronous copy.
  struct Staging {
    unsigned char* h[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<unsigned char> pageable;
    ~Staging() {
      for (int i = 0; i < 2; ++i) {
        if (ev[i]) {
          (void)hipEventSynchronize(ev[i]);
          (void)hipEventDestroy(ev[i]);
        }
        if (h[i] && pageable.empty()) (void)hipHostFree(h[i]);
      }
    }
  } stg;
  const size_t stg_bytes = static_cast<size_t>(blk) * R * es;
  if (dim <= blk) {
    stg.pageable.resize(stg_bytes);
    stg.h[0] = stg.h[1] = stg.pageable.data();
  } else {
    for (int i = 0; i < 2; ++i)
      FTLM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&stg.h[i]), stg_bytes, hipHostMallocDefault));
  }
  for (int i = 0; i < 2; ++i) FTLM_HIP_CHECK(hipEventCreateWithFlags(&stg.ev[i], hipEventDisableTiming));
  bool used_stg[2] = {false, false};
  int which = 0;
  for (int64_t b0 = 0; b0 < dim; b0 += blk, which ^= 1) {
    const int64_t nb = std::min(blk, dim - b0);
    on_threads(nthr, [&](int t) {
      for (int r = t; r < R; r += nthr) {
        std::complex<double>* q = part[static_cast<size_t>(r)].data();
        if (q0 != nullptr)
          std::memcpy(q, reinterpret_cast<const std::complex<double>*>(q0) + static_cast<int64_t>(r) * dim + b0,
                      static_cast<size_t>(nb) * sizeof(std::complex<double>));
        else
          gs.draw(r, nb, q);
        double s2 = ss[static_cast<size_t>(r)];
        for (int64_t i = 0; i < nb; ++i) s2 += std::norm(q[i]);
        ss[static_cast<size_t>(r)] = s2;
      }
    });
    if (used_stg[which]) FTLM_HIP_CHECK(hipEventSynchronize(stg.ev[which]));  // its previous copy is done
    unsigned char* host = stg.h[which];
    on_threads(nthr, [&](int t) {
      const int64_t i0 = nb * t / nthr, i1 = nb * (t + 1) / nthr;
      for (int64_t i = i0; i < i1; ++i)
        for (int r = 0; r < R; ++r) {
          const std::complex<double> z = part[static_cast<size_t>(r)][static_cast<size_t>(i)];
          unsigned char* dst = &host[(static_cast<size_t>(i) * R + r) * es];
          if (K.dtype == FTLM_C64) {
            const float2 v = make_float2(static_cast<float>(z.real()), static_cast<float>(z.imag()));
            std::memcpy(dst, &v, es);
          } else {
            const double2 v = make_double2(z.real(), z.imag());
            std::memcpy(dst, &v, es);
          }
        }
    });
    FTLM_HIP_CHECK(hipMemcpyAsync(static_cast<unsigned char*>(K.slot(0)) + static_cast<size_t>(b0) * R * es, host,
                                  static_cast<size_t>(nb) * R * es, hipMemcpyHostToDevice, st));
    FTLM_HIP_CHECK(hipEventRecord(stg.ev[which], st));
    used_stg[which] = true;
  }
  finish_start(K, ss, st);
}

// Start norms from the host sums ss[r] of |q_r|^2 (src/lanczos.cpp:166-175: 1 / ||q||, or e_0 when
// the norm is below 1e-18), then the recurrence state.
void finish_start(Krylov& K, const std::vector<double>& ss, hipStream_t st) {
  const int R = K.R;
  const size_t es = K.dtype == FTLM_C64 ? sizeof(float2) : sizeof(double2);
  std::vector<double> inv(static_cast<size_t>(R));
  for (int r = 0; r < R; ++r) {
    const double n0 = std::sqrt(ss[static_cast<size_t>(r)]);
    if (n0 < 1e-18) {  // src/lanczos.cpp:170-175: start from e_0 instead
      static const double2 one = make_double2(1.0, 0.0);
      static const float2 onef = make_float2(1.0f, 0.0f);
      FTLM_HIP_CHECK(hipMemcpyAsync(static_cast<unsigned char*>(K.slot(0)) + static_cast<size_t>(r) * es,
                                    K.dtype == FTLM_C64 ? static_cast<const void*>(&onef)
                                                        : static_cast<const void*>(&one),
                                    es, hipMemcpyHostToDevice, st));
      inv[r] = 1.0;
    } else {
      inv[r] = 1.0 / n0;
    }
  }
  if (K.dtype == FTLM_C64) {
    // complex64 storage: recompute the norm of the stored (rounded) vectors on the device
    K.start(st);
  } else {
    K.start_host_norm(inv, st);
  }
  FTLM_HIP_CHECK(hipStreamSynchronize(st));
}

void check_opts(const ftlm_ftlm_opts_t* o) {
  FTLM_REQUIRE(o != nullptr, FTLM_EINVAL, "null options");
  FTLM_REQUIRE(o->dtype == FTLM_C128 || o->dtype == FTLM_C64, FTLM_EINVAL, "bad dtype");
  FTLM_REQUIRE(o->start == FTLM_START_GAUSS_MT || o->start == FTLM_START_PHASE, FTLM_EINVAL, "bad start");
  FTLM_REQUIRE(o->quad == FTLM_QUAD_REFERENCE || o->quad == FTLM_QUAD_CONVERGED, FTLM_EINVAL, "bad quad");
  FTLM_REQUIRE(o->reorth >= FTLM_REORTH_NONE && o->reorth <= FTLM_REORTH_PARTIAL, FTLM_EINVAL, "bad reorth");
}


}  // namespace ftlm_hip

struct ftlm_krylov {
  ftlm_hip::Krylov K;
  int steps_done = 0;
};

using namespace ftlm_hip;

constexpr int kMaxSamplesPerBlock = 32;  // recurrences of one Krylov block (R of ftlm_hip_apply)

namespace ftlm_hip {
// every (sector, sample) recurrence in one persistent launch (batched.hip)
void lanczos_batched(const std::vector<const Sector*>& sectors, const std::vector<int>& Ms, int R, unsigned seed,
                     int start, int dtype, bool force_real, const std::vector<std::vector<double2>>* q0_host,
                     double* d_alpha, double* d_beta, int* d_used, int* d_alive, hipStream_t st);
}  // namespace ftlm_hip

namespace {

// ftlm_hip_ftlm_ritz for a list of small full sectors through one persistent launch.
void ritz_persistent(ftlm_op_t* const* ops, int n_ops, const ftlm_ftlm_opts_t* opts, double* ritz, double* weight,
                     int* used, double* alpha_out, double* beta_out, hipStream_t st) {
  const int R = std::max(1, opts->R), M = std::max(1, opts->M);
  std::vector<const Sector*> secs;
  std::vector<int> Ms, which;  // which[s] = operator index of batch sector s
  for (int io = 0; io < n_ops; ++io) {
    const ftlm_op* op = ops[io];
    const int64_t dim = op->dim;
    double* rz = ritz + static_cast<int64_t>(io) * R * M;
    double* wt = weight + static_cast<int64_t>(io) * R * M;
    int* us = used + static_cast<int64_t>(io) * R;
    std::fill(rz, rz + static_cast<int64_t>(R) * M, 0.0);
    std::fill(wt, wt + static_cast<int64_t>(R) * M, 0.0);
    std::fill(us, us + R, 0);
    if (alpha_out) std::fill(alpha_out + static_cast<int64_t>(io) * R * M, alpha_out + static_cast<int64_t>(io + 1) * R * M, 0.0);
    if (beta_out) std::fill(beta_out + static_cast<int64_t>(io) * R * M, beta_out + static_cast<int64_t>(io + 1) * R * M, 0.0);
    if (dim <= 0) continue;
    if (dim == 1) {  // a single state: ln Z = -beta H00 (src/ftlm_thermo.cpp:155-160 / 219-224)
      DevBuf<double2> x(1), y(1);
      const double2 one = make_double2(1.0, 0.0);
      x.upload(&one, 1, st);
      hv_apply(op, x.p, y.p, 1, FTLM_C128, nullptr, nullptr, st, opts->real_variant != 0);
      double2 h;
      FTLM_HIP_CHECK(hipMemcpyAsync(&h, y.p, sizeof(double2), hipMemcpyDeviceToHost, st));
      FTLM_HIP_CHECK(hipStreamSynchronize(st));
      for (int r = 0; r < R; ++r) {
        rz[static_cast<int64_t>(r) * M] = h.x;
        wt[static_cast<int64_t>(r) * M] = 1.0;
        us[r] = 1;
      }
      continue;
    }
    secs.push_back(static_cast<const Sector*>(op));
    Ms.push_back(opts->clamp_M_to_dim ? static_cast<int>(std::min<int64_t>(M, dim)) : M);
    which.push_back(io);
  }
  if (secs.empty()) return;
  const int ns = static_cast<int>(secs.size()), n_out = ns * R;
  std::vector<std::vector<double2>> q0;
  if (opts->start == FTLM_START_GAUSS_MT) {  // the reference's mt19937 Gaussian starts, per sample
    // every (sector, sample) start is its own serial stream: the threads take them in turn
    q0.resize(static_cast<size_t>(ns));
    for (int s = 0; s < ns; ++s) q0[s].resize(static_cast<size_t>(R) * secs[s]->dim);
    const int njobs = ns * R;
    const int nthr = std::max(1, std::min(njobs, gauss_threads(16)));
    on_threads(nthr, [&](int t) {
      std::vector<std::complex<double>> q;
      for (int j = t; j < njobs; j += nthr) {
        const int s = j / R, r = j - s * R;
        const int64_t dim = secs[s]->dim;
        gauss_mt(dim, opts->seed + static_cast<unsigned>(r) * 100003u, &q);
        for (int64_t i = 0; i < dim; ++i)
          q0[s][static_cast<size_t>(r) * dim + i] = make_double2(q[i].real(), q[i].imag());
      }
    });
  }
  DevBuf<double> da(static_cast<size_t>(M) * n_out), db(static_cast<size_t>(M) * n_out);
  DevBuf<int> du(static_cast<size_t>(n_out)), dal(static_cast<size_t>(n_out));
  FTLM_HIP_CHECK(hipMemsetAsync(da.p, 0, da.n * sizeof(double), st));
  FTLM_HIP_CHECK(hipMemsetAsync(db.p, 0, db.n * sizeof(double), st));
  FTLM_HIP_CHECK(hipMemsetAsync(du.p, 0, du.n * sizeof(int), st));
  FTLM_HIP_CHECK(hipMemsetAsync(dal.p, 0, dal.n * sizeof(int), st));
  lanczos_batched(secs, Ms, R, opts->seed, opts->start, opts->dtype, opts->real_variant != 0,
                  opts->start == FTLM_START_GAUSS_MT ? &q0 : nullptr, da.p, db.p, du.p, dal.p, st);
  std::vector<int> hu(static_cast<size_t>(n_out)), ha(static_cast<size_t>(n_out)), hnb(static_cast<size_t>(n_out));
  FTLM_HIP_CHECK(hipMemcpy(hu.data(), du.p, sizeof(int) * n_out, hipMemcpyDeviceToHost));
  FTLM_HIP_CHECK(hipMemcpy(ha.data(), dal.p, sizeof(int) * n_out, hipMemcpyDeviceToHost));
  for (int i = 0; i < n_out; ++i) hnb[i] = ha[i] ? hu[i] : hu[i] - 1;
  DevBuf<int> dnb(static_cast<size_t>(n_out));
  dnb.upload(hnb, st);
  DevBuf<double> dr(static_cast<size_t>(n_out) * M), dw(static_cast<size_t>(n_out) * M);
  FTLM_HIP_CHECK(hipMemsetAsync(dr.p, 0, dr.n * sizeof(double), st));
  FTLM_HIP_CHECK(hipMemsetAsync(dw.p, 0, dw.n * sizeof(double), st));
  quadrature_device(da.p, db.p, du.p, dnb.p, n_out, n_out, M, opts->quad, 0, 1e-14, 1, dr.p, dw.p, M, st);
  std::vector<double> hr(static_cast<size_t>(n_out) * M), hw(static_cast<size_t>(n_out) * M);
  FTLM_HIP_CHECK(hipMemcpy(hr.data(), dr.p, sizeof(double) * hr.size(), hipMemcpyDeviceToHost));
  FTLM_HIP_CHECK(hipMemcpy(hw.data(), dw.p, sizeof(double) * hw.size(), hipMemcpyDeviceToHost));
  std::vector<double> ha_(alpha_out ? static_cast<size_t>(M) * n_out : 0), hb_(beta_out ? static_cast<size_t>(M) * n_out : 0);
  if (alpha_out) FTLM_HIP_CHECK(hipMemcpy(ha_.data(), da.p, sizeof(double) * ha_.size(), hipMemcpyDeviceToHost));
  if (beta_out) FTLM_HIP_CHECK(hipMemcpy(hb_.data(), db.p, sizeof(double) * hb_.size(), hipMemcpyDeviceToHost));
  for (int s = 0; s < ns; ++s) {
    const int io = which[s];
    for (int r = 0; r < R; ++r) {
      const int o = s * R + r;
      const int64_t dst = (static_cast<int64_t>(io) * R + r) * M;
      std::copy(hr.begin() + static_cast<int64_t>(o) * M, hr.begin() + static_cast<int64_t>(o + 1) * M, ritz + dst);
      std::copy(hw.begin() + static_cast<int64_t>(o) * M, hw.begin() + static_cast<int64_t>(o + 1) * M, weight + dst);
      used[static_cast<int64_t>(io) * R + r] = hu[o];
      for (int k = 0; k < M; ++k) {
        if (alpha_out) alpha_out[dst + k] = ha_[static_cast<size_t>(k) * n_out + o];
        if (beta_out) beta_out[dst + k] = hb_[static_cast<size_t>(k) * n_out + o];
      }
    }
  }
}

// Momentum blocks that share one orbit graph run as momentum stacks: up to 16 blocks side by side,
// kMaxSamplesPerBlock / blocks samples of each per recurrence group, so one stack apply per step
// serves every block (the reference loops over K, benchmarks/bench_ftlm_nmu_4x4_k.cpp:801-1103).
// Slot (j, r) of a group starting at sample g0 is block j's sample g0 + r, started as the per-block
// path starts it (seed + (g0 + r) * 100003, random phases on the block's rows), and its alpha/beta
// go to that block's column of the batch arrays [M][S], io * R + g0 + r.  Blocks with fewer than
// two rows, or fewer rows than M when M is clamped to the dimension, stay on the per-block path.
void ritz_stacked(ftlm_op_t* const* ops, int n_ops, const ftlm_ftlm_opts_t* opts, const unsigned* op_seeds,
                  double* qa, double* qb, int* qn, int* qnb, int64_t S, int* used, double* alpha_out,
                  double* beta_out, std::vector<char>* done, std::vector<char>* from_krylov, hipStream_t st) {
  if (opts->reorth != FTLM_REORTH_NONE || opts->start != FTLM_START_PHASE || opts->real_variant) return;
  const int R = std::max(1, opts->R), M = std::max(1, opts->M);
  const int es = opts->dtype == FTLM_C64 ? 8 : 16;
  // AUTO stacks where the per-block launches would carry elements narrower than 128 bytes (R * es):
  // at 4x4 (8,8) a stack of 16 blocks then costs 1.8 ms (complex128) / 0.9 ms (complex64) per
  // sample and block against 5.4 / 2.7 ms per block at R = 1 / 2, while from 128-byte elements up
  // the per-block applies are as fast per sample (1.7 / 0.83 ms at R = 8 / 16).
  if (opts->path == FTLM_PATH_AUTO && R * es >= 128) return;
  std::map<const void*, std::vector<int>> by_graph;  // graph -> operator indices, in call order
  std::vector<const void*> order;
  for (int io = 0; io < n_ops; ++io) {
    if ((*done)[static_cast<size_t>(io)] || ops[io]->kind != kKBlock || ops[io]->dim < 2) continue;
    if (opts->clamp_M_to_dim && ops[io]->dim < M) continue;
    const void* g = kblock_graph(ops[io]);
    if (!by_graph.count(g)) order.push_back(g);
    by_graph[g].push_back(io);
  }
  for (const void* g : order) {
    const std::vector<int>& all = by_graph[g];
    if (all.size() < 2) continue;
    for (size_t c0 = 0; c0 < all.size(); c0 += kMaxStackBlocks) {
      const std::vector<int> ios(all.begin() + c0, all.begin() + std::min(all.size(), c0 + kMaxStackBlocks));
      const int nk = static_cast<int>(ios.size());
      if (nk < 2) break;  // a lone last block runs on the per-block path
      std::vector<const ftlm_op*> blk;
      for (int io : ios) blk.push_back(ops[io]);
      FTLM_HIP_CHECK(hipSetDevice(blk[0]->device));
      std::unique_ptr<ftlm_op> stack = kstack_build(blk);
      // 256-byte stacked elements (16 blocks: 2 complex64 or 1 complex128 samples each) were the
      // fastest per sample at 4x4 (8,8)
      const int rk_max = std::max(1, std::min(kMaxSamplesPerBlock / nk, 256 / (nk * es)));
      for (int g0 = 0; g0 < R; g0 += rk_max) {
        const int rk = std::min(rk_max, R - g0), W = nk * rk;
        Krylov K;
        K.init(stack.get(), W, M, opts->dtype, FTLM_REORTH_NONE);
        std::vector<unsigned> seeds(static_cast<size_t>(W), 0u);
        for (int j = 0; j < nk; ++j) {
          const unsigned s0 = op_seeds ? op_seeds[ios[j]] : opts->seed;
          for (int r = 0; r < rk; ++r) seeds[static_cast<size_t>(j) * rk + r] = s0 + static_cast<unsigned>(g0 + r) * 100003u;
        }
        fill_start(K, FTLM_START_PHASE, seeds, nullptr, st);
        if (opts->path == FTLM_PATH_GRAPH)
          K.run_graph(0, M, st);
        else
          K.run(0, M, st);
        std::vector<int> hu(static_cast<size_t>(W)), ha(static_cast<size_t>(W));
        FTLM_HIP_CHECK(hipMemcpyAsync(hu.data(), K.used.p, sizeof(int) * W, hipMemcpyDeviceToHost, st));
        FTLM_HIP_CHECK(hipMemcpyAsync(ha.data(), K.alive.p, sizeof(int) * W, hipMemcpyDeviceToHost, st));
        for (int j = 0; j < nk; ++j) {  // slots (j, 0..rk) -> columns io_j * R + g0 ..
          const int64_t s0 = static_cast<int64_t>(ios[j]) * R + g0;
          FTLM_HIP_CHECK(hipMemcpy2DAsync(qa + s0, sizeof(double) * S, K.alpha.p + static_cast<int64_t>(j) * rk,
                                          sizeof(double) * W, sizeof(double) * rk, M, hipMemcpyDeviceToDevice, st));
          FTLM_HIP_CHECK(hipMemcpy2DAsync(qb + s0, sizeof(double) * S, K.beta.p + static_cast<int64_t>(j) * rk,
                                          sizeof(double) * W, sizeof(double) * rk, M, hipMemcpyDeviceToDevice, st));
        }
        FTLM_HIP_CHECK(hipStreamSynchronize(st));
        std::vector<double> a, b;
        if (alpha_out || beta_out) {
          a.resize(static_cast<size_t>(M) * W);
          b.resize(static_cast<size_t>(M) * W);
          FTLM_HIP_CHECK(hipMemcpy(a.data(), K.alpha.p, a.size() * sizeof(double), hipMemcpyDeviceToHost));
          FTLM_HIP_CHECK(hipMemcpy(b.data(), K.beta.p, b.size() * sizeof(double), hipMemcpyDeviceToHost));
        }
        for (int j = 0; j < nk; ++j) {
          const int io = ios[j];
          const int64_t s0 = static_cast<int64_t>(io) * R + g0;
          std::vector<int> hnb(static_cast<size_t>(rk));
          for (int r = 0; r < rk; ++r) {
            const int v = j * rk + r;
            hnb[r] = ha[v] ? hu[v] : hu[v] - 1;
            used[s0 + r] = hu[v];
            (*from_krylov)[static_cast<size_t>(s0 + r)] = 1;
            for (int k = 0; k < M; ++k) {
              const int64_t o = (s0 + r) * M + k;
              if (alpha_out) alpha_out[o] = a[static_cast<size_t>(k) * W + v];
              if (beta_out) beta_out[o] = b[static_cast<size_t>(k) * W + v];
            }
          }
          FTLM_HIP_CHECK(hipMemcpyAsync(qn + s0, hu.data() + j * rk, sizeof(int) * rk, hipMemcpyHostToDevice, st));
          FTLM_HIP_CHECK(hipMemcpyAsync(qnb + s0, hnb.data(), sizeof(int) * rk, hipMemcpyHostToDevice, st));
          FTLM_HIP_CHECK(hipStreamSynchronize(st));  // hnb is reused for the next block
        }
      }
      for (int io : ios) (*done)[static_cast<size_t>(io)] = 1;
    }
  }
}

constexpr int64_t kPersistentMaxDim = 262144;     // FTLM_PATH_PERSISTENT
constexpr int64_t kPersistentAutoMaxDim = 16384;  // FTLM_PATH_AUTO

bool persistent_eligible(ftlm_op_t* const* ops, int n_ops, int path, int reorth) {
  if (path == FTLM_PATH_PER_OP || path == FTLM_PATH_GRAPH) return false;
  if (reorth != FTLM_REORTH_NONE) {
    FTLM_REQUIRE(path != FTLM_PATH_PERSISTENT, FTLM_EUNSUPPORTED,
                 "the persistent path keeps three vectors per recurrence and cannot reorthogonalise");
    return false;
  }
  const int64_t lim = path == FTLM_PATH_PERSISTENT ? kPersistentMaxDim : kPersistentAutoMaxDim;
  for (int i = 0; i < n_ops; ++i) {
    if (ops[i] == nullptr || ops[i]->kind != kSector || ops[i]->dim > lim) {
      FTLM_REQUIRE(path != FTLM_PATH_PERSISTENT, FTLM_EUNSUPPORTED,
                   "the persistent path takes full sectors of at most 262144 states");
      return false;
    }
  }
  return n_ops > 0;
}

}  // namespace


extern "C" {

int ftlm_hip_krylov_create(const ftlm_op_t* op, int R, int max_steps, int dtype, int start, unsigned seed,
                           ftlm_krylov_t** out) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && out != nullptr, FTLM_EINVAL, "null argument");
    *out = nullptr;
    FTLM_REQUIRE(R >= 1 && R <= 32 && max_steps >= 1, FTLM_EINVAL, "R in [1,32], max_steps >= 1");
    FTLM_REQUIRE(dtype == FTLM_C128 || dtype == FTLM_C64, FTLM_EINVAL, "bad dtype");
    FTLM_REQUIRE(start == FTLM_START_GAUSS_MT || start == FTLM_START_PHASE, FTLM_EINVAL, "bad start");
    FTLM_HIP_CHECK(hipSetDevice(op->device));
    auto s = std::make_unique<ftlm_krylov>();
    s->K.init(op, R, max_steps, dtype);
    std::vector<unsigned> seeds(static_cast<size_t>(R));
    // a momentum stack's slot (j, r) starts as block j's sample r
    const int per = op->kind == kKStack ? R / std::max(1, kstack_blocks(op)) : R;
    for (int r = 0; r < R; ++r) seeds[r] = seed + static_cast<unsigned>(r % std::max(1, per)) * 100003u;
    fill_start(s->K, start, seeds, nullptr, nullptr);
    *out = s.release();
  });
}

int ftlm_hip_krylov_run(ftlm_krylov_t* k, int n_steps, void* stream) {
  return guarded([&] {
    FTLM_REQUIRE(k != nullptr && n_steps >= 0, FTLM_EINVAL, "bad argument");
    const int k1 = std::min(k->K.M, k->steps_done + n_steps);
    k->K.run(k->steps_done, k1, static_cast<hipStream_t>(stream));
    k->steps_done = k1;
  });
}

int ftlm_hip_krylov_run_graph(ftlm_krylov_t* k, int n_steps, void* stream) {
  return guarded([&] {
    FTLM_REQUIRE(k != nullptr && n_steps >= 0, FTLM_EINVAL, "bad argument");
    FTLM_HIP_CHECK(hipSetDevice(k->K.op->device));
    const int k1 = std::min(k->K.M, k->steps_done + n_steps);
    k->K.run_graph(k->steps_done, k1, static_cast<hipStream_t>(stream));
    k->steps_done = k1;
  });
}

int ftlm_hip_krylov_result(ftlm_krylov_t* k, double* alpha, double* beta, int* used, int* alive) {
  return guarded([&] {
    FTLM_REQUIRE(k != nullptr, FTLM_EINVAL, "null session");
    FTLM_HIP_CHECK(hipDeviceSynchronize());
    const size_t n = static_cast<size_t>(k->K.M) * k->K.R;
    if (alpha) FTLM_HIP_CHECK(hipMemcpy(alpha, k->K.alpha.p, n * sizeof(double), hipMemcpyDeviceToHost));
    if (beta) FTLM_HIP_CHECK(hipMemcpy(beta, k->K.beta.p, n * sizeof(double), hipMemcpyDeviceToHost));
    if (used) FTLM_HIP_CHECK(hipMemcpy(used, k->K.used.p, k->K.R * sizeof(int), hipMemcpyDeviceToHost));
    if (alive) FTLM_HIP_CHECK(hipMemcpy(alive, k->K.alive.p, k->K.R * sizeof(int), hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_krylov_bytes(const ftlm_krylov_t* k, int64_t* bytes) {
  return guarded([&] {
    FTLM_REQUIRE(k != nullptr && bytes != nullptr, FTLM_EINVAL, "null argument");
    *bytes = static_cast<int64_t>(3 * k->K.vbytes + k->K.dotp.n * sizeof(double2) + k->K.normp.n * sizeof(double));
  });
}

int ftlm_hip_krylov_destroy(ftlm_krylov_t* k) {
  return guarded([&] { delete k; });
}

int ftlm_hip_lanczos(const ftlm_op_t* op, int max_steps, unsigned seed, const double* q0, double* alpha, double* beta,
                     int* n_beta, int* used) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && alpha != nullptr && beta != nullptr && n_beta && used, FTLM_EINVAL, "null argument");
    *used = 0;
    *n_beta = 0;
    if (op->dim <= 0 || max_steps <= 0) return;
    FTLM_HIP_CHECK(hipSetDevice(op->device));
    Krylov K;
    K.init(op, 1, max_steps, FTLM_C128);
    fill_start(K, FTLM_START_GAUSS_MT, {seed}, q0, nullptr);
    K.run(0, max_steps, nullptr);
    int u = 0;
    FTLM_HIP_CHECK(hipMemcpy(&u, K.used.p, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<double> a(static_cast<size_t>(max_steps)), b(static_cast<size_t>(max_steps));
    FTLM_HIP_CHECK(hipMemcpy(a.data(), K.alpha.p, sizeof(double) * max_steps, hipMemcpyDeviceToHost));
    FTLM_HIP_CHECK(hipMemcpy(b.data(), K.beta.p, sizeof(double) * max_steps, hipMemcpyDeviceToHost));
    int alive = 0;
    FTLM_HIP_CHECK(hipMemcpy(&alive, K.alive.p, sizeof(int), hipMemcpyDeviceToHost));
    *used = u;
    *n_beta = alive ? u : u - 1;  // a breakdown step records alpha but not beta
    std::copy(a.begin(), a.begin() + u, alpha);
    std::copy(b.begin(), b.begin() + *n_beta, beta);
  });
}

int ftlm_hip_lanczos_extrema(const ftlm_op_t* op, int max_steps, unsigned seed, double* min_eval, double* max_eval,
                             int* steps_used) {
  std::vector<double> a(static_cast<size_t>(std::max(max_steps, 1))), b(static_cast<size_t>(std::max(max_steps, 1)));
  int nb = 0, u = 0;
  int rc = ftlm_hip_lanczos(op, max_steps, seed, nullptr, a.data(), b.data(), &nb, &u);
  if (rc != FTLM_OK) return rc;
  if (min_eval) *min_eval = 0.0;
  if (max_eval) *max_eval = 0.0;
  if (steps_used) *steps_used = u;
  if (u == 0) return FTLM_OK;
  double lo = 0.0, hi = 0.0;
  rc = ftlm_hip_tridiagonal_extrema(a.data(), u, b.data(), u >= 2 ? u - 1 : 0, &lo, &hi);
  if (rc != FTLM_OK) return rc;
  if (min_eval) *min_eval = lo;
  if (max_eval) *max_eval = hi;
  return FTLM_OK;
}

// ftlm_hip_ftlm_ritz with an optional seed per operator (op_seeds[io] in place of opts->seed).
static void ritz_impl(ftlm_op_t* const* ops, int n_ops, const ftlm_ftlm_opts_t* opts0, const unsigned* op_seeds, double* ritz,
               double* weight, int* used, double* alpha_out, double* beta_out, void* stream) {
  check_opts(opts0);
  FTLM_REQUIRE(ops != nullptr && n_ops >= 1 && ritz && weight && used, FTLM_EINVAL, "null argument");
  FTLM_REQUIRE(opts0->path >= FTLM_PATH_AUTO && opts0->path <= FTLM_PATH_STACK, FTLM_EINVAL, "bad path");
  ftlm_ftlm_opts_t opts_one = *opts0;
  const ftlm_ftlm_opts_t* opts = &opts_one;
  // one seed for every operator: the plain call (the persistent launch draws from opts->seed)
  bool one_seed = true;
  if (op_seeds) {
    for (int io = 1; io < n_ops; ++io) one_seed = one_seed && op_seeds[io] == op_seeds[0];
    if (one_seed) {
      opts_one.seed = op_seeds[0];
      op_seeds = nullptr;
    }
  }
  {
    const int R = std::max(1, opts->R), M = std::max(1, opts->M);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (one_seed && persistent_eligible(ops, n_ops, opts->path, opts->reorth)) {
      FTLM_HIP_CHECK(hipSetDevice(ops[0]->device));
      ritz_persistent(ops, n_ops, opts, ritz, weight, used, alpha_out, beta_out, st);
      return;
    }
    for (int io = 0; io < n_ops; ++io)
      FTLM_REQUIRE(ops[io] != nullptr && ops[io]->device == ops[0]->device, FTLM_EINVAL,
                   "the operators of one call must live on one device");
    FTLM_HIP_CHECK(hipSetDevice(ops[0]->device));
    // AUTO over a mix of sizes: the small full sectors (at least two) share one persistent launch,
    // whose results are scattered to their operators' slots; the rest run operator by operator.
    const int64_t per = static_cast<int64_t>(R) * M;
    std::vector<char> done(static_cast<size_t>(n_ops), 0);
    if (one_seed && opts->path == FTLM_PATH_AUTO && opts->reorth == FTLM_REORTH_NONE) {
      std::vector<ftlm_op_t*> small;
      std::vector<int> at;
      for (int io = 0; io < n_ops; ++io)
        if (ops[io]->kind == kSector && ops[io]->dim <= kPersistentAutoMaxDim) {
          small.push_back(ops[io]);
          at.push_back(io);
        }
      if (small.size() >= 2) {
        const size_t ns = small.size();
        std::vector<double> sr(ns * per), sw(ns * per), sa(alpha_out ? ns * per : 0), sb(beta_out ? ns * per : 0);
        std::vector<int> su(ns * R);
        ritz_persistent(small.data(), static_cast<int>(ns), opts, sr.data(), sw.data(), su.data(),
                        alpha_out ? sa.data() : nullptr, beta_out ? sb.data() : nullptr, st);
        for (size_t j = 0; j < ns; ++j) {
          const int64_t io = at[j], src = static_cast<int64_t>(j) * per, dst = io * per;
          std::copy(sr.begin() + src, sr.begin() + src + per, ritz + dst);
          std::copy(sw.begin() + src, sw.begin() + src + per, weight + dst);
          std::copy(su.begin() + static_cast<int64_t>(j) * R, su.begin() + static_cast<int64_t>(j + 1) * R,
                    used + io * R);
          if (alpha_out) std::copy(sa.begin() + src, sa.begin() + src + per, alpha_out + dst);
          if (beta_out) std::copy(sb.begin() + src, sb.begin() + src + per, beta_out + dst);
          done[static_cast<size_t>(io)] = 1;
        }
      }
    }
    // The recurrences run operator by operator; their tridiagonals are gathered into [M][S] device
    // arrays (S = n_ops * R samples) and diagonalised together in one launch after the loop (a
    // per-operator launch leaves the GPU idle: R workgroups each).
    const int64_t S = static_cast<int64_t>(n_ops) * R;
    DevBuf<double> qa(static_cast<size_t>(M) * S), qb(static_cast<size_t>(M) * S);
    DevBuf<int> qn(static_cast<size_t>(S)), qnb(static_cast<size_t>(S));
    FTLM_HIP_CHECK(hipMemsetAsync(qa.p, 0, qa.n * sizeof(double), st));
    FTLM_HIP_CHECK(hipMemsetAsync(qb.p, 0, qb.n * sizeof(double), st));
    FTLM_HIP_CHECK(hipMemsetAsync(qn.p, 0, qn.n * sizeof(int), st));
    FTLM_HIP_CHECK(hipMemsetAsync(qnb.p, 0, qnb.n * sizeof(int), st));
    std::vector<char> from_krylov(static_cast<size_t>(S), 0);
    if (opts->path == FTLM_PATH_AUTO || opts->path == FTLM_PATH_STACK)
      ritz_stacked(ops, n_ops, opts, op_seeds, qa.p, qb.p, qn.p, qnb.p, S, used, alpha_out, beta_out, &done,
                   &from_krylov, st);
    for (int io = 0; io < n_ops; ++io) {
      if (done[static_cast<size_t>(io)]) continue;
      const ftlm_op* op0 = ops[io];
      FTLM_REQUIRE(op0 != nullptr, FTLM_EINVAL, "null operator");
      FTLM_HIP_CHECK(hipSetDevice(op0->device));
      const ftlm_op* op = op0;
      const int64_t dim = op->dim;
      double* rz = ritz + static_cast<int64_t>(io) * R * M;
      double* wt = weight + static_cast<int64_t>(io) * R * M;
      int* us = used + static_cast<int64_t>(io) * R;
      std::fill(rz, rz + static_cast<int64_t>(R) * M, 0.0);
      std::fill(wt, wt + static_cast<int64_t>(R) * M, 0.0);
      if (dim <= 0) {
        std::fill(us, us + R, 0);
        continue;
      }
      int Mop = M;
      if (opts->clamp_M_to_dim) Mop = static_cast<int>(std::min<int64_t>(M, dim));
      if (dim == 1) {
        // src/ftlm_thermo.cpp:155-160 / 219-224: a single state, lnZ = -beta H00
        DevBuf<double2> x(1), y(1);
        const double2 one = make_double2(1.0, 0.0);
        x.upload(&one, 1, st);
        hv_apply(op, x.p, y.p, 1, FTLM_C128, nullptr, nullptr, st, opts->real_variant != 0);
        double2 h;
        FTLM_HIP_CHECK(hipMemcpyAsync(&h, y.p, sizeof(double2), hipMemcpyDeviceToHost, st));
        FTLM_HIP_CHECK(hipStreamSynchronize(st));
        for (int r = 0; r < R; ++r) {
          rz[static_cast<int64_t>(r) * M] = h.x;
          wt[static_cast<int64_t>(r) * M] = 1.0;
          us[r] = 1;
        }
        continue;
      }
      // samples run in groups of at most kMaxSamplesPerBlock recurrences; sample r keeps its seed
      for (int g0 = 0; g0 < R; g0 += kMaxSamplesPerBlock) {
        const int Rg = std::min(kMaxSamplesPerBlock, R - g0);
        Krylov K;
        K.init(op, Rg, Mop, opts->dtype, opts->reorth);
        K.force_real = opts->real_variant != 0;
        std::vector<unsigned> seeds(static_cast<size_t>(Rg));
        const unsigned seed0 = op_seeds ? op_seeds[io] : opts->seed;
        for (int r = 0; r < Rg; ++r) seeds[r] = seed0 + static_cast<unsigned>(g0 + r) * 100003u;
        fill_start(K, opts->start, seeds, nullptr, st);
        if (opts->path == FTLM_PATH_GRAPH && op->kind != kCallback && op->kind != kGram)
          K.run_graph(0, Mop, st);
        else
          K.run(0, Mop, st);
        // this group's alpha/beta [Mop][Rg] into columns s0 .. s0+Rg of the batch arrays
        const int64_t s0 = static_cast<int64_t>(io) * R + g0;
        std::vector<int> hu(static_cast<size_t>(Rg)), ha(static_cast<size_t>(Rg));
        FTLM_HIP_CHECK(hipMemcpyAsync(hu.data(), K.used.p, sizeof(int) * Rg, hipMemcpyDeviceToHost, st));
        FTLM_HIP_CHECK(hipMemcpyAsync(ha.data(), K.alive.p, sizeof(int) * Rg, hipMemcpyDeviceToHost, st));
        FTLM_HIP_CHECK(hipMemcpy2DAsync(qa.p + s0, sizeof(double) * S, K.alpha.p, sizeof(double) * Rg,
                                        sizeof(double) * Rg, Mop, hipMemcpyDeviceToDevice, st));
        FTLM_HIP_CHECK(hipMemcpy2DAsync(qb.p + s0, sizeof(double) * S, K.beta.p, sizeof(double) * Rg,
                                        sizeof(double) * Rg, Mop, hipMemcpyDeviceToDevice, st));
        FTLM_HIP_CHECK(hipStreamSynchronize(st));
        std::vector<int> hnb(static_cast<size_t>(Rg));
        for (int r = 0; r < Rg; ++r) hnb[r] = ha[r] ? hu[r] : hu[r] - 1;
        FTLM_HIP_CHECK(hipMemcpyAsync(qn.p + s0, hu.data(), sizeof(int) * Rg, hipMemcpyHostToDevice, st));
        FTLM_HIP_CHECK(hipMemcpyAsync(qnb.p + s0, hnb.data(), sizeof(int) * Rg, hipMemcpyHostToDevice, st));
        std::copy(hu.begin(), hu.end(), us + g0);
        std::fill(from_krylov.begin() + s0, from_krylov.begin() + s0 + Rg, 1);
        if (alpha_out || beta_out) {
          std::vector<double> a(static_cast<size_t>(Mop) * Rg), b(static_cast<size_t>(Mop) * Rg);
          FTLM_HIP_CHECK(hipMemcpyAsync(a.data(), K.alpha.p, a.size() * sizeof(double), hipMemcpyDeviceToHost, st));
          FTLM_HIP_CHECK(hipMemcpyAsync(b.data(), K.beta.p, b.size() * sizeof(double), hipMemcpyDeviceToHost, st));
          FTLM_HIP_CHECK(hipStreamSynchronize(st));
          for (int r = 0; r < Rg; ++r)
            for (int k = 0; k < M; ++k) {
              const int64_t o = (static_cast<int64_t>(io) * R + g0 + r) * M + k;
              if (alpha_out) alpha_out[o] = k < Mop ? a[static_cast<size_t>(k) * Rg + r] : 0.0;
              if (beta_out) beta_out[o] = k < Mop ? b[static_cast<size_t>(k) * Rg + r] : 0.0;
            }
        }
        FTLM_HIP_CHECK(hipStreamSynchronize(st));  // the block's buffers are freed at scope end
      }
    }
    if (std::find(from_krylov.begin(), from_krylov.end(), 1) != from_krylov.end()) {
      DevBuf<double> dr(static_cast<size_t>(S) * M), dw(static_cast<size_t>(S) * M);
      FTLM_HIP_CHECK(hipMemsetAsync(dr.p, 0, dr.n * sizeof(double), st));
      FTLM_HIP_CHECK(hipMemsetAsync(dw.p, 0, dw.n * sizeof(double), st));
      quadrature_device(qa.p, qb.p, qn.p, qnb.p, static_cast<int>(S), static_cast<int>(S), M, opts->quad, 0, 1e-14,
                        1, dr.p, dw.p, M, st);
      std::vector<double> hr(static_cast<size_t>(S) * M), hw(static_cast<size_t>(S) * M);
      FTLM_HIP_CHECK(hipMemcpy(hr.data(), dr.p, sizeof(double) * hr.size(), hipMemcpyDeviceToHost));
      FTLM_HIP_CHECK(hipMemcpy(hw.data(), dw.p, sizeof(double) * hw.size(), hipMemcpyDeviceToHost));
      for (int64_t sm = 0; sm < S; ++sm)
        if (from_krylov[static_cast<size_t>(sm)]) {
          std::copy(hr.begin() + sm * M, hr.begin() + (sm + 1) * M, ritz + sm * M);
          std::copy(hw.begin() + sm * M, hw.begin() + (sm + 1) * M, weight + sm * M);
        }
    }
  }
}

int ftlm_hip_ftlm_ritz(ftlm_op_t* const* ops, int n_ops, const ftlm_ftlm_opts_t* opts, double* ritz, double* weight,
                       int* used, double* alpha_out, double* beta_out, void* stream) {
  return guarded([&] { ritz_impl(ops, n_ops, opts, nullptr, ritz, weight, used, alpha_out, beta_out, stream); });
}

int ftlm_hip_ftlm_ritz_seeded(ftlm_op_t* const* ops, int n_ops, const ftlm_ftlm_opts_t* opts, const unsigned* op_seeds,
                              double* ritz, double* weight, int* used, double* alpha_out, double* beta_out,
                              void* stream) {
  return guarded([&] {
    FTLM_REQUIRE(op_seeds != nullptr, FTLM_EINVAL, "null seed list");
    ritz_impl(ops, n_ops, opts, op_seeds, ritz, weight, used, alpha_out, beta_out, stream);
  });
}

int ftlm_hip_ftlm_log_partition(const ftlm_op_t* op, const double* betas, int nb, const ftlm_ftlm_opts_t* opts,
                                double* lnZ) {
  return guarded([&] {
    check_opts(opts);
    FTLM_REQUIRE(op != nullptr && betas != nullptr && lnZ != nullptr && nb >= 1, FTLM_EINVAL, "null argument");
    const int R = std::max(1, opts->R), M = std::max(1, opts->M);
    if (op->dim <= 0) {
      for (int i = 0; i < nb; ++i) lnZ[i] = -INFINITY;
      return;
    }
    std::vector<double> rz(static_cast<size_t>(R) * M), wt(static_cast<size_t>(R) * M);
    std::vector<int> us(static_cast<size_t>(R));
    ftlm_op_t* ops[1] = {const_cast<ftlm_op_t*>(op)};
    int rc = ftlm_hip_ftlm_ritz(ops, 1, opts, rz.data(), wt.data(), us.data(), nullptr, nullptr, nullptr);
    if (rc != FTLM_OK) throw Error(rc, ftlm_hip_last_error());
    const int64_t dim = op->dim;
    rc = ftlm_hip_ftlm_thermo(rz.data(), wt.data(), us.data(), 1, R, M, &dim, nullptr, betas, nb, nullptr, 0, lnZ,
                              nullptr, nullptr, nullptr, nullptr);
    if (rc != FTLM_OK) throw Error(rc, ftlm_hip_last_error());
    if (dim == 1) {
      // the reference returns -beta * H00 without the ln(dim) term (ln 1 = 0 anyway)
      for (int i = 0; i < nb; ++i) lnZ[i] = -betas[i] * rz[0];
    }
  });
}

}  // extern "C"
