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
//
// Two kernels:
//  k_hv_rows  (D_dn >= 128): a workgroup owns a tile of 256 consecutive ids of ONE row iu, so
//             everything that depends on u is wavefront-uniform: the up-bond loop runs on scalar
//             loads, and the down-bond coefficients are pre-signed with the row's cross-spin JW
//             parity into an LDS table once per row.  Down-bond entries are read slot-major
//             (one contiguous 256 B run per wavefront and slot).  Tiles are dealt to the 8 XCDs
//             in contiguous row ranges so that the blocks gathering inside a row share one L2.
//  k_hv_flat  (small sectors): one thread per flattened element, no row uniformity assumed.

#include <cstdlib>

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
constexpr int kHvMaxGrid = 8192;  // bounds the dot-partial array
constexpr int kXcd = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin

// acc += c * v for RB vectors (real or complex coefficient).
template <typename T, int RB, bool REAL>
__device__ __forceinline__ void axpy_block(T (&ar)[RB], T (&ai)[RB], const typename Cx<T>::v2* __restrict__ src,
                                           T cr, T ci) {
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    const typename Cx<T>::v2 v = src[r];
    if (REAL) {
      ar[r] += cr * v.x;
      ai[r] += cr * v.y;
    } else {
      ar[r] += cr * v.x - ci * v.y;
      ai[r] += cr * v.y + ci * v.x;
    }
  }
}

// Store y = xs * acc and accumulate <xs x_own | y> (FP64).
template <typename T, int RB>
__device__ __forceinline__ void finish_element(const T (&ar)[RB], const T (&ai)[RB], const typename Cx<T>::v2 (&xo)[RB],
                                               typename Cx<T>::v2* __restrict__ ye, const double* __restrict__ xs,
                                               int r0, bool want_dot, double (&dre)[RB], double (&dim)[RB]) {
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    T yr = ar[r], yi = ai[r];
    double sx = 1.0;
    if (xs) {
      sx = xs[r0 + r];
      yr = static_cast<T>(static_cast<double>(yr) * sx);
      yi = static_cast<T>(static_cast<double>(yi) * sx);
    }
    typename Cx<T>::v2 o;
    o.x = yr;
    o.y = yi;
    ye[r] = o;
    if (want_dot) {
      const double qr = sx * static_cast<double>(xo[r].x), qi = sx * static_cast<double>(xo[r].y);
      dre[r] += qr * static_cast<double>(yr) + qi * static_cast<double>(yi);
      dim[r] += qr * static_cast<double>(yi) - qi * static_cast<double>(yr);
    }
  }
}

template <int RB>
__device__ __forceinline__ void block_dot_store(const double (&dre)[RB], const double (&dim)[RB], double2* dotp, int RT,
                                                int r0) {
  __shared__ double red[kHvBlock / 64][RB][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    const double a = wave_sum(dre[r]), b = wave_sum(dim[r]);
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

__device__ __forceinline__ double diag_value(const SectorView& S, uint32_t u, uint32_t d, int iu, int id) {
  const uint4 pl = S.planes[iu];
  const int cross = __popc(d & pl.x) + 2 * __popc(d & pl.y) + 4 * __popc(d & pl.z) + 8 * __popc(d & pl.w);
  return S.U * static_cast<double>(__popc(u & d)) + S.V * static_cast<double>(S.nuu[iu] + S.ndd[id] + cross);
}

// ------------------------------------------------------------------ row-tiled kernel
template <typename T, int RB, bool REAL>
__global__ __launch_bounds__(kHvBlock) void k_hv_rows(const SectorView S, const typename Cx<T>::v2* __restrict__ x,
                                                      typename Cx<T>::v2* __restrict__ y, const int RT, const int r0,
                                                      const double* __restrict__ xs, double2* __restrict__ dotp,
                                                      const int nchunk, const int64_t ntiles) {
  using V2 = typename Cx<T>::v2;
  __shared__ T s_cr[kMaxBondsLds], s_ci[kMaxBondsLds];
  double dre[RB], dimg[RB];
#pragma unroll
  for (int r = 0; r < RB; ++r) dre[r] = dimg[r] = 0.0;

  const int xcd = blockIdx.x % kXcd, lb = blockIdx.x / kXcd, nbx = gridDim.x / kXcd;
  const int64_t per = (ntiles + kXcd - 1) / kXcd;
  const int64_t t_begin = xcd * per, t_end = min(ntiles, t_begin + per);
  int cur_row = -1;
  const int Ddn = S.Ddn;
  for (int64_t tile = t_begin + lb; tile < t_end; tile += nbx) {
    const int iu = static_cast<int>(tile / nchunk);
    const int chunk = static_cast<int>(tile - static_cast<int64_t>(iu) * nchunk);
    const uint32_t u = S.up_mask[iu];
    if (iu != cur_row) {  // uniform: every thread of the block sees the same tile
      __syncthreads();
      for (int b = threadIdx.x; b < S.nbond_dn; b += kHvBlock) {
        const double2 c = S.coef_dn[b];
        const bool neg = (__popc(u & S.xmask_dn[b]) & 1u) != 0;
        s_cr[b] = static_cast<T>(neg ? -c.x : c.x);
        s_ci[b] = static_cast<T>(neg ? -c.y : c.y);
      }
      __syncthreads();
      cur_row = iu;
    }
    const int id = chunk * kHvBlock + static_cast<int>(threadIdx.x);
    if (id >= Ddn) continue;
    const uint32_t d = S.dn_mask[id];
    const double dg = diag_value(S, u, d, iu, id);
    const int64_t e = static_cast<int64_t>(iu) * Ddn + id;

    T ar[RB], ai[RB];
    V2 xo[RB];
    const V2* xe = x + e * RT + r0;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      xo[r] = xe[r];
      ar[r] = static_cast<T>(dg) * xo[r].x;
      ai[r] = static_cast<T>(dg) * xo[r].y;
    }

    // up-spin bonds: the entry list, bond data and source row are wavefront-uniform
    const int ku0 = S.off_up[iu], ku1 = S.off_up[iu + 1];
    for (int k = ku0; k < ku1; ++k) {
      const uint32_t ent = S.ent_up[k];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const uint32_t par = (ent >> 31) ^ (__popc(d & S.xmask_up[b]) & 1u);
      const double2 c = S.coef_up[b];
      const T cr = static_cast<T>(par ? -c.x : c.x);
      const T ci = static_cast<T>(par ? -c.y : c.y);
      axpy_block<T, RB, REAL>(ar, ai, x + (static_cast<int64_t>(t) * Ddn + id) * RT + r0, cr, ci);
    }
    // down-spin bonds: per-lane entries, slot-major table, row-signed coefficients in LDS
    const int kd = S.cnt_dn[id];
    const int64_t rowbase = static_cast<int64_t>(iu) * Ddn;
    for (int k = 0; k < kd; ++k) {
      const uint32_t ent = S.entT_dn[static_cast<int64_t>(k) * Ddn + id];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const bool neg = (ent >> 31) != 0;
      const T cr = neg ? -s_cr[b] : s_cr[b];
      const T ci = neg ? -s_ci[b] : s_ci[b];
      axpy_block<T, RB, REAL>(ar, ai, x + (rowbase + t) * RT + r0, cr, ci);
    }
    finish_element<T, RB>(ar, ai, xo, y + e * RT + r0, xs, r0, dotp != nullptr, dre, dimg);
  }
  if (dotp) block_dot_store<RB>(dre, dimg, dotp, RT, r0);
}

// ------------------------------------------------------------------ flat kernel (small sectors)
template <typename T, int RB, bool REAL>
__global__ __launch_bounds__(kHvBlock) void k_hv_flat(const SectorView S, const typename Cx<T>::v2* __restrict__ x,
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
    const int iu = static_cast<int>(e / S.Ddn);
    const int id = static_cast<int>(e - static_cast<int64_t>(iu) * S.Ddn);
    const uint32_t u = S.up_mask[iu], d = S.dn_mask[id];
    const double dg = diag_value(S, u, d, iu, id);
    T ar[RB], ai[RB];
    V2 xo[RB];
    const V2* xe = x + e * RT + r0;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      xo[r] = xe[r];
      ar[r] = static_cast<T>(dg) * xo[r].x;
      ai[r] = static_cast<T>(dg) * xo[r].y;
    }
    for (int k = S.off_up[iu]; k < S.off_up[iu + 1]; ++k) {
      const uint32_t ent = S.ent_up[k];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const uint32_t par = (ent >> 31) ^ (__popc(d & S.xmask_up[b]) & 1u);
      const double2 c = S.coef_up[b];
      axpy_block<T, RB, REAL>(ar, ai, x + (static_cast<int64_t>(t) * S.Ddn + id) * RT + r0,
                              static_cast<T>(par ? -c.x : c.x), static_cast<T>(par ? -c.y : c.y));
    }
    const int64_t rowbase = static_cast<int64_t>(iu) * S.Ddn;
    for (int k = S.off_dn[id]; k < S.off_dn[id + 1]; ++k) {
      const uint32_t ent = S.ent_dn[k];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const uint32_t par = (ent >> 31) ^ (__popc(u & S.xmask_dn[b]) & 1u);
      const double2 c = S.coef_dn[b];
      axpy_block<T, RB, REAL>(ar, ai, x + (rowbase + t) * RT + r0, static_cast<T>(par ? -c.x : c.x),
                              static_cast<T>(par ? -c.y : c.y));
    }
    finish_element<T, RB>(ar, ai, xo, y + e * RT + r0, xs, r0, dotp != nullptr, dre, dimg);
  }
  if (dotp) block_dot_store<RB>(dre, dimg, dotp, RT, r0);
}

namespace {

bool use_rows(const SectorView& S) { return S.Ddn >= 128; }

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : dflt;
}

// Rows kernel: blocks per XCD sized so that the rows being gathered concurrently on one XCD
// (about bpx tiles of 256 elements x RT vectors) stay near the 4 MB L2 of that XCD.
int grid_size(const SectorView& S, int RT, int elem_bytes) {
  if (use_rows(S)) {
    const int nchunk = (S.Ddn + kHvBlock - 1) / kHvBlock;
    const int64_t ntiles = static_cast<int64_t>(S.Dup) * nchunk;
    const int64_t per = (ntiles + kXcd - 1) / kXcd;
    const int64_t window = static_cast<int64_t>(env_int("FTLM_HV_WINDOW_KB", 4096)) * 1024;
    int64_t bpx = window / (static_cast<int64_t>(kHvBlock) * RT * elem_bytes);
    bpx = std::max<int64_t>(64, std::min<int64_t>(bpx, kHvMaxGrid / kXcd));
    bpx = env_int("FTLM_HV_BPX", static_cast<int>(bpx));
    bpx = std::max<int64_t>(1, std::min<int64_t>(bpx, kHvMaxGrid / kXcd));
    return static_cast<int>(kXcd * std::min<int64_t>(per, bpx));
  }
  const int64_t tiles = (S.dim + kHvBlock - 1) / kHvBlock;
  return static_cast<int>(std::min<int64_t>(tiles, kHvMaxGrid));
}

template <typename T, int RB, bool REAL>
void launch(const SectorView& S, const void* x, void* y, int RT, int r0, const double* xs, double2* dotp,
            hipStream_t st) {
  using V2 = typename Cx<T>::v2;
  const dim3 g(grid_size(S, RT, sizeof(V2))), b(kHvBlock);
  const V2* xp = static_cast<const V2*>(x);
  V2* yp = static_cast<V2*>(y);
  if (use_rows(S)) {
    const int nchunk = (S.Ddn + kHvBlock - 1) / kHvBlock;
    const int64_t ntiles = static_cast<int64_t>(S.Dup) * nchunk;
    hipLaunchKernelGGL((k_hv_rows<T, RB, REAL>), g, b, 0, st, S, xp, yp, RT, r0, xs, dotp, nchunk, ntiles);
  } else {
    hipLaunchKernelGGL((k_hv_flat<T, RB, REAL>), g, b, 0, st, S, xp, yp, RT, r0, xs, dotp);
  }
}

template <typename T, bool REAL>
void launch_chunk(const SectorView& S, const void* x, void* y, int RT, int r0, int rb, const double* xs,
                  double2* dotp, hipStream_t st) {
  switch (rb) {
    case 16: launch<T, 16, REAL>(S, x, y, RT, r0, xs, dotp, st); break;
    case 8: launch<T, 8, REAL>(S, x, y, RT, r0, xs, dotp, st); break;
    case 4: launch<T, 4, REAL>(S, x, y, RT, r0, xs, dotp, st); break;
    case 2: launch<T, 2, REAL>(S, x, y, RT, r0, xs, dotp, st); break;
    default: launch<T, 1, REAL>(S, x, y, RT, r0, xs, dotp, st); break;
  }
}

}  // namespace

// Upper bound of the partial-sum blocks any H*V launch on this operator writes.
int hv_partial_blocks(const ftlm_op* op) {
  (void)op;
  return kHvMaxGrid;
}

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
  return grid_size(v, R, dtype == FTLM_C64 ? sizeof(float2) : sizeof(double2));
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
