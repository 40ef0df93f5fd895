// hv.hip -- matrix-free extended-Hubbard H * V in a full (N_up, N_dn) sector (gfx950).
//
// Replaces apply_extended_hubbard (src/hubbard_hamiltonian.cpp:51-108).  The reference scatters
// ("push") every column into hashed rows; here every output element gathers ("pull") its ~68
// sources, so no atomics and no hashing are needed:
//
//   y[iu,id] = diag(u,d) x[iu,id]
//            + sum_{up bonds b into u}   s * c_b * x[iu', id]   (coalesced: the whole wavefront
//                                                               reads one contiguous row segment)
//            + sum_{down bonds b into d} s * c_b * x[iu, id']   (gather inside row iu)
//
// with the interleaved Jordan-Wigner sign s = (-1)^(own-spin parity from the table
//                                                 + popcount(other string & bond mask)).
// The sign depends only on modes strictly between the bond ends, so evaluating it on the row
// (target) state equals the reference's evaluation on the column state.
//
// diag(u,d) = U popcount(u&d) + V (sum_pairs u_i u_j + sum_pairs d_i d_j + sum_s d_s nb_u(s)),
// the last term from four bit planes of nb_u (number of occupied pair partners of s).
//
// A launch handles RB vectors [r0, r0+RB) of a block of RT vectors laid out v[elem*RT + r];
// decode work (tables, signs, addresses) is shared by the RB vectors.

#include "common.hpp"

namespace ftlm_hip {

template <typename T>
struct Cx;
template <>
struct Cx<double> {
  using v2 = double2;
};
template <>
struct Cx<float> {
  using v2 = float2;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kHvBlock = 256;
constexpr int kHvMaxGrid = 8192;  // persistent-style grid: bounds the dot-partial array

template <typename T, int RB, bool REAL>
__global__ __launch_bounds__(kHvBlock) void k_hv_sector(const SectorView S, const typename Cx<T>::v2* __restrict__ x,
                                                        typename Cx<T>::v2* __restrict__ y, const int RT, const int r0,
                                                        const double* __restrict__ xs, double2* __restrict__ dotp) {
  using V2 = typename Cx<T>::v2;
  double dre[RB], dimg[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) dre[r] = dimg[r] = 0.0;

  const int64_t ntiles = (S.dim + kHvBlock - 1) / kHvBlock;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t e = tile * kHvBlock + threadIdx.x;
    if (e >= S.dim) continue;
    const uint32_t iu = static_cast<uint32_t>(static_cast<uint64_t>(e) / static_cast<uint32_t>(S.Ddn));
    const uint32_t id = static_cast<uint32_t>(e - static_cast<int64_t>(iu) * S.Ddn);
    const uint32_t u = S.up_mask[iu], d = S.dn_mask[id];

    const uint4 pl = S.planes[iu];
    const int cross = __popc(d & pl.x) + 2 * __popc(d & pl.y) + 4 * __popc(d & pl.z) + 8 * __popc(d & pl.w);
    const double dg = S.U * static_cast<double>(__popc(u & d)) +
                      S.V * static_cast<double>(S.nuu[iu] + S.ndd[id] + cross);

    T ar[RB], ai[RB];
    V2 xo[RB];
    const V2* xe = x + static_cast<int64_t>(e) * RT + r0;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      xo[r] = xe[r];
      ar[r] = static_cast<T>(dg) * xo[r].x;
      ai[r] = static_cast<T>(dg) * xo[r].y;
    }

    // up-spin bonds: source row iu' = target of the entry, same column id
    const int ku1 = S.off_up[iu + 1];
    for (int k = S.off_up[iu]; k < ku1; ++k) {
      const uint32_t ent = S.ent_up[k];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const uint32_t par = (ent >> 31) ^ (__popc(d & S.xmask_up[b]) & 1u);
      const double2 c = S.coef_up[b];
      const T cr = static_cast<T>(par ? -c.x : c.x);
      const V2* src = x + (static_cast<int64_t>(t) * S.Ddn + id) * RT + r0;
      if (REAL) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const V2 v = src[r];
          ar[r] += cr * v.x;
          ai[r] += cr * v.y;
        }
      } else {
        const T ci = static_cast<T>(par ? -c.y : c.y);
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const V2 v = src[r];
          ar[r] += cr * v.x - ci * v.y;
          ai[r] += cr * v.y + ci * v.x;
        }
      }
    }
    // down-spin bonds: source column id' = target of the entry, same row iu
    const int kd1 = S.off_dn[id + 1];
    const int64_t rowbase = static_cast<int64_t>(iu) * S.Ddn;
    for (int k = S.off_dn[id]; k < kd1; ++k) {
      const uint32_t ent = S.ent_dn[k];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const uint32_t par = (ent >> 31) ^ (__popc(u & S.xmask_dn[b]) & 1u);
      const double2 c = S.coef_dn[b];
      const T cr = static_cast<T>(par ? -c.x : c.x);
      const V2* src = x + (rowbase + t) * RT + r0;
      if (REAL) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const V2 v = src[r];
          ar[r] += cr * v.x;
          ai[r] += cr * v.y;
        }
      } else {
        const T ci = static_cast<T>(par ? -c.y : c.y);
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const V2 v = src[r];
          ar[r] += cr * v.x - ci * v.y;
          ai[r] += cr * v.y + ci * v.x;
        }
      }
    }

    V2* ye = y + static_cast<int64_t>(e) * RT + r0;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      T yr = ar[r], yi = ai[r];
      double sx = 1.0;
      if (xs) {
        sx = xs[r0 + r];
        yr = static_cast<T>(static_cast<double>(yr) * sx);
        yi = static_cast<T>(static_cast<double>(yi) * sx);
      }
      V2 o;
      o.x = yr;
      o.y = yi;
      ye[r] = o;
      if (dotp) {  // <xs x | y> in FP64
        const double qr = sx * static_cast<double>(xo[r].x), qi = sx * static_cast<double>(xo[r].y);
        dre[r] += qr * static_cast<double>(yr) + qi * static_cast<double>(yi);
        dimg[r] += qr * static_cast<double>(yi) - qi * static_cast<double>(yr);
      }
    }
  }

  if (dotp) {
    __shared__ double red[kHvBlock / 64][RB][2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const double a = wave_sum(dre[r]), b = wave_sum(dimg[r]);
      if (lane == 0) {
        red[wv][r][0] = a;
        red[wv][r][1] = b;
      }
    }
    __syncthreads();
    if (threadIdx.x < RB) {
      double a = 0.0, b = 0.0;
      for (int w = 0; w < kHvBlock / 64; ++w) {
        a += red[w][threadIdx.x][0];
        b += red[w][threadIdx.x][1];
      }
      dotp[static_cast<int64_t>(blockIdx.x) * RT + r0 + threadIdx.x] = make_double2(a, b);
    }
  }
}

namespace {

int grid_size(int64_t dim) {
  const int64_t tiles = (dim + kHvBlock - 1) / kHvBlock;
  return static_cast<int>(std::min<int64_t>(tiles, kHvMaxGrid));
}

template <typename T, bool REAL>
void launch_chunk(const SectorView& S, const void* x, void* y, int RT, int r0, int rb, const double* xs,
                  double2* dotp, hipStream_t st) {
  using V2 = typename Cx<T>::v2;
  const dim3 g(grid_size(S.dim)), b(kHvBlock);
  const V2* xp = static_cast<const V2*>(x);
  V2* yp = static_cast<V2*>(y);
  switch (rb) {
    case 16: hipLaunchKernelGGL((k_hv_sector<T, 16, REAL>), g, b, 0, st, S, xp, yp, RT, r0, xs, dotp); break;
    case 8: hipLaunchKernelGGL((k_hv_sector<T, 8, REAL>), g, b, 0, st, S, xp, yp, RT, r0, xs, dotp); break;
    case 4: hipLaunchKernelGGL((k_hv_sector<T, 4, REAL>), g, b, 0, st, S, xp, yp, RT, r0, xs, dotp); break;
    case 2: hipLaunchKernelGGL((k_hv_sector<T, 2, REAL>), g, b, 0, st, S, xp, yp, RT, r0, xs, dotp); break;
    default: hipLaunchKernelGGL((k_hv_sector<T, 1, REAL>), g, b, 0, st, S, xp, yp, RT, r0, xs, dotp); break;
  }
}

}  // namespace

int hv_partial_blocks(const ftlm_op* op) { return grid_size(op->dim); }

// Block apply over all RT vectors in chunks of RB (<= 8 for complex128, <= 16 for complex64).
int hv_apply_sector(const Sector* S, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st, bool force_real) {
  const bool real = S->real_h || force_real;
  const SectorView v = S->view();
  const int maxrb = dtype == FTLM_C64 ? 16 : 8;
  for (int r0 = 0; r0 < R;) {
    int rb = 1;
    while (rb * 2 <= maxrb && r0 + rb * 2 <= R) rb *= 2;
    if (dtype == FTLM_C64) {
      if (real) launch_chunk<float, true>(v, x, y, R, r0, rb, xs, dotp, st);
      else launch_chunk<float, false>(v, x, y, R, r0, rb, xs, dotp, st);
    } else {
      if (real) launch_chunk<double, true>(v, x, y, R, r0, rb, xs, dotp, st);
      else launch_chunk<double, false>(v, x, y, R, r0, rb, xs, dotp, st);
    }
    r0 += rb;
  }
  check_launch("k_hv_sector");
  return grid_size(S->dim);
}

int hv_apply_kblock(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st);

int hv_apply(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
             hipStream_t st, bool force_real) {
  if (op->kind == kSector)
    return hv_apply_sector(static_cast<const Sector*>(op), x, y, R, dtype, xs, dotp, st, force_real);
  return hv_apply_kblock(op, x, y, R, dtype, xs, dotp, st);
}

}  // namespace ftlm_hip

using namespace ftlm_hip;

extern "C" {

int ftlm_hip_apply(const ftlm_op_t* op, const void* x, void* y, int R, int dtype, void* stream) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && x != nullptr && y != nullptr, FTLM_EINVAL, "null argument");
    FTLM_REQUIRE(R >= 1 && R <= 32, FTLM_EINVAL, "R must be in [1, 32]");
    FTLM_REQUIRE(dtype == FTLM_C128 || dtype == FTLM_C64, FTLM_EINVAL, "dtype must be FTLM_C128 or FTLM_C64");
    FTLM_REQUIRE(x != y, FTLM_EINVAL, "x and y must not alias");
    hv_apply(op, x, y, R, dtype, nullptr, nullptr, static_cast<hipStream_t>(stream));
  });
}

int ftlm_hip_apply_host(const ftlm_op_t* op, const double* x, double* y) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && x != nullptr && y != nullptr, FTLM_EINVAL, "null argument");
    FTLM_HIP_CHECK(hipSetDevice(op->device));
    const size_t n = static_cast<size_t>(op->dim);
    DevBuf<double2> dx(n), dy(n);
    dx.upload(reinterpret_cast<const double2*>(x), n);
    hv_apply(op, dx.p, dy.p, 1, FTLM_C128, nullptr, nullptr, nullptr);
    FTLM_HIP_CHECK(hipMemcpy(y, dy.p, n * sizeof(double2), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
