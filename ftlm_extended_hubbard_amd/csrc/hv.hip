// hv.hip -- matrix-free extended-Hubbard H * V in a full (N_up, N_dn) sector (gfx950).
//
// Replaces apply_extended_hubbard (src/hubbard_hamiltonian.cpp:51-108).  The reference scatters
// ("push") every column into hashed rows; here every output element gathers ("pull") its ~68
// sources, so no atomics and no hashing are needed:
//
//   y[iu,id] = diag(u,d) x[iu,id]
//            + sum_{down bonds b into d} s * c_b * x[iu, id']   (sources inside row iu)
//            + sum_{up bonds b into u}   s * c_b * x[iu', id]   (sources inside column id)
//
// with the interleaved Jordan-Wigner sign s = (-1)^(own-spin parity from the table
//                                                 + popcount(other string & bond mask)).
// The sign depends only on modes strictly between the bond ends, so evaluating it on the row
// (target) state equals the reference's evaluation on the column state.
// diag(u,d) = U popcount(u&d) + V (sum_pairs u_i u_j + sum_pairs d_i d_j + sum_s d_s nb_u(s)),
// the last term from four bit planes of nb_u (number of occupied pair partners of s).
//
// Cache structure.  Every source element is read by ~34 outputs of its row (down bonds) and ~34
// of its column (up bonds).  One row of R vectors is D_dn * R * sizeof(complex) bytes (206 KB for
// one complex128 vector, 1.65 MB for 16 complex64 vectors), so an XCD's 4 MB L2 can hold the
// few rows it is working on but not the ~34 other rows an up bond reaches.  Large sectors
// therefore use two passes:
//   k_hv_dn  down bonds, tiles = (row, chunk of ids) in row-major order.  u is uniform per
//            tile: the down-bond coefficients are pre-signed with the row's cross-spin parity
//            in LDS.
//   k_hv_up  up bonds and the diagonal, y += ... ; tiles = (stripe of W adjacent ids, chunk
//            of rows) in stripe-major order; a stripe of all rows (D_up * W * R * s) is kept
//            below ~3 MB.
// Workgroups are dealt to the 8 XCDs round-robin (blockIdx % 8).  Each XCD owns a contiguous
// share of the tiles and hands them out through an atomic counter, so the tiles in flight on an
// XCD always form one contiguous window (a few rows, or one stripe) that its L2 holds, no
// matter how unevenly its workgroups progress.  Hop tables are slot-major, so the table slice
// of the window is contiguous too.
// The R vectors of an element are spread over L lanes of 16 bytes, so a wavefront load reads
// whole 128-byte lines.  Per-hop addressing is a 32-bit offset from a per-row (per-column) base
// pointer; complex updates are four explicit FMAs.  Small sectors (D_dn < 128) use one flat
// single-pass kernel.

#include <array>
#include <cstdlib>

#include <cstring>
#include <thread>

#include "common.hpp"

namespace ftlm_hip {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kHvBlock = 256;
constexpr int kHvMaxGrid = 8192;   // total dot-partial blocks of one apply
constexpr int kPassMaxGrid = 4096; // per pass
constexpr int kXcd = 8;            // MI355X: 8 XCDs, workgroups dealt round-robin

// Lane layout of an element's RB vectors: L lanes of 16 bytes (VPL vectors each).
template <typename T, int RB>
struct Split {
  static constexpr int BYTES = RB * static_cast<int>(sizeof(typename Cx<T>::v2));
  static constexpr int L = BYTES >= 32 ? BYTES / 16 : 1;
  static constexpr int VPL = RB / L;
  static constexpr int EPB = kHvBlock / L;  // elements per pass of a workgroup
  static_assert(RB % L == 0 && kHvBlock % L == 0, "bad vector split");
};

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c) {
  return fma(a, b, c);
}

// acc += (cr + i ci) * v for N vectors.
template <typename T, int N, bool REAL>
__device__ __forceinline__ void axpy_block(T (&ar)[N], T (&ai)[N], const typename Cx<T>::v2* __restrict__ src, T cr,
                                           T ci) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const typename Cx<T>::v2 v = src[r];
    ar[r] = fmaT(cr, v.x, ar[r]);
    ai[r] = fmaT(cr, v.y, ai[r]);
    if (!REAL) {
      ar[r] = fmaT(-ci, v.y, ar[r]);
      ai[r] = fmaT(ci, v.x, ai[r]);
    }
  }
}

// Summation order of the two-pass H*V kernels (gather and step passes), so that all of them give
// the same bits:
//   y = (part(S_down) + part(S_up)) + part(T(diag) * x),   part(a) = T(double(a) * xs)
// where S_down and S_up are the down- and up-bond hop sums, each accumulated from 0 in increasing
// source order, and the two additions are done in T.  The row pass stores part(S_down); the
// column pass adds the rest (it reads x at its targets anyway for the <xs x | y> partials).
// Contraction is off so that no kernel fuses a product into a sum.  The one-pass kernel of small
// sectors (k_hv_flat) keeps the reference's order, diagonal first: y = part(diag x + S_down) +
// part(S_up), so that the late, rounding-sensitive Lanczos coefficients of the small sectors the
// reference's drivers print follow the reference's own.
template <typename T>
__device__ __forceinline__ T hv_part(T a, double sx) {
#pragma clang fp contract(off)
  return static_cast<T>(static_cast<double>(a) * sx);
}
template <typename T>
__device__ __forceinline__ T hv_finish(T row_part, T s_up, T dg, T xq, double sx) {
#pragma clang fp contract(off)
  const T uu = static_cast<T>(static_cast<double>(s_up) * sx);
  const T ug = static_cast<T>(static_cast<double>(dg * xq) * sx);
  return (row_part + uu) + ug;
}

// acc += (cr + i ci) * v for N vectors already in registers.
template <typename T, int N, bool REAL>
__device__ __forceinline__ void axpy_regs(T (&ar)[N], T (&ai)[N], const typename Cx<T>::v2 (&v)[N], T cr, T ci) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    ar[r] = fmaT(cr, v[r].x, ar[r]);
    ai[r] = fmaT(cr, v[r].y, ai[r]);
    if (!REAL) {
      ar[r] = fmaT(-ci, v[r].y, ar[r]);
      ai[r] = fmaT(ci, v[r].x, ai[r]);
    }
  }
}

// Streaming accesses of y (written by the row pass, read and rewritten by the column pass) can
// carry the non-temporal hint, so that they do not displace the gathered source window in L2.
// Build knob, off by default.
#ifndef FTLM_HV_NT
#define FTLM_HV_NT 0
#endif
typedef float ftlm_f2v __attribute__((ext_vector_type(2)));
typedef double ftlm_d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void st_stream(float2* p, float2 v) {
#if FTLM_HV_NT
  ftlm_f2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<ftlm_f2v*>(p));
#else
  *p = v;
#endif
}
__device__ __forceinline__ void st_stream(double2* p, double2 v) {
#if FTLM_HV_NT
  ftlm_d2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<ftlm_d2v*>(p));
#else
  *p = v;
#endif
}
__device__ __forceinline__ float2 ld_stream(const float2* p) {
#if FTLM_HV_NT
  const ftlm_f2v w = __builtin_nontemporal_load(reinterpret_cast<const ftlm_f2v*>(p));
  return make_float2(w.x, w.y);
#else
  return *p;
#endif
}
__device__ __forceinline__ double2 ld_stream(const double2* p) {
#if FTLM_HV_NT
  const ftlm_d2v w = __builtin_nontemporal_load(reinterpret_cast<const ftlm_d2v*>(p));
  return make_double2(w.x, w.y);
#else
  return *p;
#endif
}

#ifndef FTLM_HV_UNROLL
#define FTLM_HV_UNROLL 8
#endif
constexpr int kUnroll = FTLM_HV_UNROLL;  // hops whose gathers are issued together (divides kSlotPad)

// Hop entries of a tile are staged in LDS (16-byte loads, then LDS reads shared by the L lanes
// of an element) when a tile is at most 64 elements wide; wider single-vector tiles read them
// straight from the L2-resident table, which costs fewer vector-memory instructions.
template <int EPB>
struct StageEntries {
  static constexpr bool value = EPB <= 64;
};

// Block reduction of per-lane dot partials; lane t holds vectors [r0 + (t % L) * VPL, ... + VPL).
template <int L, int VPL, int RB>
__device__ __forceinline__ void store_dot(const double (&dre)[VPL], const double (&dim)[VPL], double2* dotp,
                                          int64_t slot, int RT, int r0) {
  __shared__ double red[kHvBlock][VPL][2];
  __syncthreads();
#pragma unroll
  for (int v = 0; v < VPL; ++v) {
    red[threadIdx.x][v][0] = dre[v];
    red[threadIdx.x][v][1] = dim[v];
  }
  __syncthreads();
  if (threadIdx.x < RB) {
    const int s2 = threadIdx.x / VPL, v2 = threadIdx.x % VPL;
    double a = 0.0, b = 0.0;
    for (int j = s2; j < kHvBlock; j += L) {
      a += red[j][v2][0];
      b += red[j][v2][1];
    }
    dotp[slot * RT + r0 + threadIdx.x] = make_double2(a, b);
  }
}

// Diagonal of element (iu, id) from the strings u, d, the bit planes of row iu and the pair
// counts nuu[iu], ndd[id].
__device__ __forceinline__ double diag_from(const SectorView& S, uint32_t u, uint32_t d, uint4 pl, int nu, int nd) {
  const int cross = __popc(d & pl.x) + 2 * __popc(d & pl.y) + 4 * __popc(d & pl.z) + 8 * __popc(d & pl.w);
  return S.U * static_cast<double>(__popc(u & d)) + S.V * static_cast<double>(nu + nd + cross);
}

__device__ __forceinline__ double diag_value(const SectorView& S, uint32_t u, uint32_t d, int iu, int id) {
  return diag_from(S, u, d, S.planes[iu], S.nuu[iu], S.ndd[id]);
}

// ------------------------------------------------------------------ per-XCD tile queues
// Each pass owns 9 counters: [0..7] next tile of each XCD's share, [8] finished workgroups.  A
// workgroup (on XCD blockIdx % 8) repeatedly takes the next tile of its XCD's contiguous share,
// so the tiles in flight on one XCD always form one contiguous window whatever the relative
// speed of its workgroups.  The last workgroup to finish returns the counters to zero.
__device__ __forceinline__ int next_tile(int* ctr, int xcd, int* s_tile) {
  __syncthreads();
  if (threadIdx.x == 0) *s_tile = atomicAdd(&ctr[xcd], 1);
  __syncthreads();
  return *s_tile;
}

__device__ __forceinline__ void queue_done(int* ctr) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&ctr[kXcd], 1) == static_cast<int>(gridDim.x) - 1) {
      for (int i = 0; i < kXcd; ++i) atomicExch(&ctr[i], 0);
      atomicExch(&ctr[kXcd], 0);
    }
  }
}

// ------------------------------------------------------------------ rows pass: down bonds
// Tiles: (row iu, chunk of EPB ids), row-major; XCD x owns tiles [x*per, (x+1)*per).
// Per hop: one entry read, one 24-bit multiply-add for the source offset from the (uniform) row
// base, one LDS read of the pre-signed coefficient (index = entry >> 20), and the complex FMAs.
// Hop lists are padded to multiples of kUnroll with zero-coefficient entries, so a group of
// kUnroll gathers is issued without bounds tests.
template <typename T, int RB, bool REAL>
__global__ __launch_bounds__(kHvBlock) void k_hv_dn(const SectorView S, const typename Cx<T>::v2* __restrict__ x,
                                                    typename Cx<T>::v2* __restrict__ y, const int RT, const int r0,
                                                    const double* __restrict__ xs,
                                                    int* __restrict__ ctr, const int nchunk, const int ntiles) {
  using V2 = typename Cx<T>::v2;
  using SP = Split<T, RB>;
  constexpr int L = SP::L, VPL = SP::VPL, EPB = SP::EPB;
  constexpr bool kStage = StageEntries<EPB>::value;
  static_assert(kSlotPad % kUnroll == 0, "hop groups must divide the slot padding");
  __shared__ V2 s_c[kCoefSlots];  // row-signed coefficients [parity * 512 + bond]
  __shared__ int s_tile;
  extern __shared__ uint4 s_ent4[];  // this tile's entries, [maxcnt][EPB] uint32
  const uint32_t* s_ent = reinterpret_cast<const uint32_t*>(s_ent4);
  const int xcd = blockIdx.x % kXcd;
  const int sub = static_cast<int>(threadIdx.x) % L, el = static_cast<int>(threadIdx.x) / L;
  const int rr = r0 + sub * VPL;
  const int per = (ntiles + kXcd - 1) / kXcd;
  const int t0 = min(ntiles, xcd * per), nmine = min(ntiles, t0 + per) - t0;
  const int Ddn = S.Ddn;
  const int64_t row_stride = static_cast<int64_t>(Ddn) * RT;
  const uint32_t es = static_cast<uint32_t>(RT * sizeof(V2));      // bytes per element
  const uint32_t lane_b = static_cast<uint32_t>(rr * sizeof(V2));  // this lane's bytes in an element
  const int mc = S.maxcnt_dn, estride = S.entT_stride_dn;
  double sxv[VPL];
#pragma unroll
  for (int v = 0; v < VPL; ++v) sxv[v] = xs ? xs[rr + v] : 1.0;
  int cur_row = -1, cur_chunk = -1;
  for (int local = next_tile(ctr, xcd, &s_tile); local < nmine; local = next_tile(ctr, xcd, &s_tile)) {
    const int tile = __builtin_amdgcn_readfirstlane(t0 + local);
    const int iu = tile / nchunk;
    const int chunk = tile - iu * nchunk;
    const uint32_t u = S.line_umask ? S.line_umask[iu] : S.up_mask[iu];
    // the row's first element (caller-described lines: line_base[iu])
    const int64_t rowoff = S.line_base ? S.line_base[iu] * RT : static_cast<int64_t>(iu) * row_stride;
    // uniform branches: the whole block works on one tile
    if (kStage && chunk != cur_chunk) {  // the chunk's entries: one 16-byte load per 4 strings and slot
      const uint4* src4 = reinterpret_cast<const uint4*>(S.entT_dn + chunk * EPB);
      for (int i = threadIdx.x; i < mc * (EPB / 4); i += kHvBlock) {
        const int k = i / (EPB / 4), q = i - k * (EPB / 4);
        s_ent4[i] = src4[(static_cast<int64_t>(k) * estride) / 4 + q];
      }
      cur_chunk = chunk;
    }
    if (iu != cur_row) {
      for (int b = threadIdx.x; b < kMaxBondsLds; b += kHvBlock) {
        V2 c;
        c.x = c.y = static_cast<T>(0);
        if (b < S.nbond_dn) {
          const double2 cc = S.coef_dn[b];
          const bool neg = (__popc(u & S.xmask_dn[b]) & 1u) != 0;
          c.x = static_cast<T>(neg ? -cc.x : cc.x);
          c.y = static_cast<T>(neg ? -cc.y : cc.y);
        }
        s_c[b] = c;
        c.x = -c.x;
        c.y = -c.y;
        s_c[b + kMaxBondsLds] = c;
      }
      cur_row = iu;
    }
    __syncthreads();
    const int id = chunk * EPB + el;
    if (id >= Ddn) continue;
    const char* xb = reinterpret_cast<const char*>(x + rowoff);  // uniform
    T ar[VPL], ai[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) ar[v] = ai[v] = static_cast<T>(0);
    const int kd = S.cnt_dn[id];
    // staged: slot k of this element at s_ent[k * EPB + el]; otherwise straight from the table
    const uint32_t* ent = kStage ? s_ent + el : S.entT_dn + id;
    const int eld = kStage ? EPB : estride;
    // kUnroll sources are loaded before any is consumed (independent gathers in flight);
    // accumulation stays in entry order.
    for (int k0 = 0; k0 < kd; k0 += kUnroll) {
      uint32_t en[kUnroll];
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) en[j] = ent[(k0 + j) * eld];
      V2 src[kUnroll][VPL];
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) {
        const V2* sp = reinterpret_cast<const V2*>(xb + (__umul24(en[j] & kTargetMask, es) + lane_b));
#pragma unroll
        for (int v = 0; v < VPL; ++v) src[j][v] = sp[v];
      }
#pragma unroll
      for (int j = 0; j < kUnroll; ++j) {
        const V2 c = s_c[en[j] >> kTargetBits];
        axpy_regs<T, VPL, REAL>(ar, ai, src[j], c.x, c.y);
      }
    }
    V2* ye = reinterpret_cast<V2*>(reinterpret_cast<char*>(y + rowoff) +
                                   (__umul24(static_cast<uint32_t>(id), es) + lane_b));
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
      V2 o;
      o.x = hv_part(ar[v], sxv[v]);
      o.y = hv_part(ai[v], sxv[v]);
      st_stream(&ye[v], o);
    }
  }
  queue_done(ctr);
}

// Up-bond accumulation of one element over its padded hop list.  ONE_ID: every lane of the
// block works on the same id (stripes one id wide), so the cross-spin sign popcount(d & mask_b)
// is already folded into s_c; otherwise each lane applies it from s_xm.
template <typename T, int VPL, bool REAL, bool ONE_ID>
__device__ __forceinline__ void up_hops(T (&ar)[VPL], T (&ai)[VPL], const typename Cx<T>::v2* __restrict__ xcol,
                                        const uint32_t* __restrict__ ent, int eld, int ku, uint32_t rs24,
                                        const typename Cx<T>::v2* s_c, const uint32_t* s_xm, uint32_t d) {
  using V2 = typename Cx<T>::v2;
  for (int k0 = 0; k0 < ku; k0 += kUnroll) {
    uint32_t en[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) en[j] = ent[(k0 + j) * eld];
    V2 src[kUnroll][VPL];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      // 20-bit row index times a 24-bit row stride: a 48-bit element offset (mul_u24 + mul_hi_u24)
      const uint64_t off = static_cast<uint64_t>(en[j] & kTargetMask) * static_cast<uint64_t>(rs24 & 0xffffffu);
      const V2* sp = xcol + off;
#pragma unroll
      for (int v = 0; v < VPL; ++v) src[j][v] = sp[v];
    }
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      const uint32_t ci = en[j] >> kTargetBits;
      V2 c = s_c[ci];
      if (!ONE_ID && (__popc(d & s_xm[ci & kBondMask]) & 1u)) {
        c.x = -c.x;
        c.y = -c.y;
      }
      axpy_regs<T, VPL, REAL>(ar, ai, src[j], c.x, c.y);
    }
  }
}

// ------------------------------------------------------------------ column pass: up bonds + diagonal, y +=
// Tiles: (stripe of W ids, chunk of EPB/W rows), stripe-major; element e of a tile ->
// (row chunk*EPB/W + e/W, id stripe*W + e%W).  While an XCD works through one stripe (all rows x
// W ids <= ~3 MB) every up bond of the stripe reads from that XCD's L2.
template <typename T, int RB, bool REAL>
__global__ __launch_bounds__(kHvBlock) void k_hv_up(const SectorView S, const typename Cx<T>::v2* __restrict__ x,
                                                    typename Cx<T>::v2* __restrict__ y, const int RT, const int r0,
                                                    const double* __restrict__ xs, double2* __restrict__ dotp,
                                                    const int64_t slot0, int* __restrict__ ctr, const int wlog,
                                                    const int tps, const int ntiles) {
  using V2 = typename Cx<T>::v2;
  using SP = Split<T, RB>;
  constexpr int L = SP::L, VPL = SP::VPL, EPB = SP::EPB;
  constexpr bool kStage = StageEntries<EPB>::value;
  __shared__ V2 s_c[kCoefSlots];  // [parity * 512 + bond]; signed by d when stripes are one id wide
  __shared__ uint32_t s_xm[kMaxBondsLds];
  __shared__ int s_tile;
  const bool one_id = wlog == 0;
  auto fill_coef = [&](uint32_t dm) {
    for (int b = threadIdx.x; b < kMaxBondsLds; b += kHvBlock) {
      V2 c;
      c.x = c.y = static_cast<T>(0);
      uint32_t xm = 0;
      if (b < S.nbond_up) {
        const double2 cc = S.coef_up[b];
        xm = S.xmask_up[b];
        const bool neg = (__popc(dm & xm) & 1u) != 0;
        c.x = static_cast<T>(neg ? -cc.x : cc.x);
        c.y = static_cast<T>(neg ? -cc.y : cc.y);
      }
      s_c[b] = c;
      c.x = -c.x;
      c.y = -c.y;
      s_c[b + kMaxBondsLds] = c;
      s_xm[b] = xm;
    }
  };
  if (!one_id) fill_coef(0u);
  const int xcd = blockIdx.x % kXcd;
  const int sub = static_cast<int>(threadIdx.x) % L, el = static_cast<int>(threadIdx.x) / L;
  const int rr = r0 + sub * VPL;
  const int W = 1 << wlog;
  const int rows_per_tile = EPB >> wlog;
  const int per = (ntiles + kXcd - 1) / kXcd;
  const int t0 = min(ntiles, xcd * per), nmine = min(ntiles, t0 + per) - t0;
  const int Ddn = S.Ddn, Dup = S.Dup;
  const int64_t row_stride = static_cast<int64_t>(Ddn) * RT;
  const uint32_t rs24 = static_cast<uint32_t>(row_stride);  // < 2^24, checked by the host
  double sxv[VPL], dre[VPL], dimg[VPL];
#pragma unroll
  for (int v = 0; v < VPL; ++v) {
    sxv[v] = xs ? xs[rr + v] : 1.0;
    dre[v] = dimg[v] = 0.0;
  }
  extern __shared__ uint4 s_ent4[];  // this tile's entries, [maxcnt][rows_per_tile] uint32
  uint32_t* s_ent = reinterpret_cast<uint32_t*>(s_ent4);
  const int mc = S.maxcnt_up, estride = S.entT_stride_up;
  int cur_rc = -1, cur_stripe = -1;
  for (int local = next_tile(ctr, xcd, &s_tile); local < nmine; local = next_tile(ctr, xcd, &s_tile)) {
    const int tile = __builtin_amdgcn_readfirstlane(t0 + local);
    const int stripe = tile / tps;
    const int rc = tile - stripe * tps;
    bool changed = false;
    if (kStage && rc != cur_rc) {  // uniform: stage the row chunk's entries
      const int rbase = rc * rows_per_tile;
      if ((rows_per_tile & 3) == 0) {
        const int q4 = rows_per_tile / 4;
        const uint4* src4 = reinterpret_cast<const uint4*>(S.entT_up + rbase);
        for (int i = threadIdx.x; i < mc * q4; i += kHvBlock) {
          const int k = i / q4, q = i - k * q4;
          s_ent4[i] = src4[(static_cast<int64_t>(k) * estride) / 4 + q];
        }
      } else {
        for (int i = threadIdx.x; i < mc * rows_per_tile; i += kHvBlock) {
          const int k = i / rows_per_tile, q = i - k * rows_per_tile;
          s_ent[i] = S.entT_up[static_cast<int64_t>(k) * estride + rbase + q];
        }
      }
      cur_rc = rc;
      changed = true;
    }
    if (one_id && stripe != cur_stripe) {  // uniform: sign the coefficients by this id's string
      fill_coef(S.dn_mask[stripe]);
      cur_stripe = stripe;
      changed = true;
    }
    if (changed) __syncthreads();
    const int id = (stripe << wlog) + (el & (W - 1));
    const int rl = el >> wlog;
    const int iu = rc * rows_per_tile + rl;
    if (id >= Ddn || iu >= Dup) continue;
    const uint32_t d = S.dn_mask[id];
    const V2* xcol = x + static_cast<int64_t>(id) * RT + rr;
    T ar[VPL], ai[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) ar[v] = ai[v] = static_cast<T>(0);
    const int ku = S.cnt_up[iu];
    const uint32_t* ent = kStage ? s_ent + rl : S.entT_up + iu;
    const int eld = kStage ? rows_per_tile : estride;
    if (one_id)
      up_hops<T, VPL, REAL, true>(ar, ai, xcol, ent, eld, ku, rs24, s_c, s_xm, d);
    else
      up_hops<T, VPL, REAL, false>(ar, ai, xcol, ent, eld, ku, rs24, s_c, s_xm, d);
    const int64_t eo = iu * row_stride + static_cast<int64_t>(id) * RT + rr;
    V2* ye = y + eo;
    const T dgt = static_cast<T>(diag_value(S, S.up_mask[iu], d, iu, id));
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
      const V2 prev = ld_stream(&ye[v]);
      const V2 q = x[eo + v];
      V2 o;
      o.x = hv_finish(prev.x, ar[v], dgt, q.x, sxv[v]);
      o.y = hv_finish(prev.y, ai[v], dgt, q.y, sxv[v]);
      st_stream(&ye[v], o);
      if (dotp) {  // <xs x | y> partials of the finished y
        const double qr = sxv[v] * static_cast<double>(q.x), qi = sxv[v] * static_cast<double>(q.y);
        dre[v] += qr * static_cast<double>(o.x) + qi * static_cast<double>(o.y);
        dimg[v] += qr * static_cast<double>(o.y) - qi * static_cast<double>(o.x);
      }
    }
  }
  queue_done(ctr);
  if (dotp) store_dot<L, VPL, RB>(dre, dimg, dotp, slot0 + blockIdx.x, RT, r0);
}

// ------------------------------------------------------------------ LDS-staged step passes
// Elements of 128 bytes (16 complex64 or 8 complex128 vectors) use the step tables (layout at
// kStepBlock in common.hpp).  A workgroup of kStepThreads threads takes tiles (line, block of
// kStepBlock targets) from its XCD's queue and keeps the block's sums in registers (64 per
// lane: kStepRounds rounds x 32 bytes of an element, as complex components).  For every step of the
// block, the step's source elements and each wavefront's entry rows are copied into LDS by
// asynchronous global->LDS loads one step ahead (two buffers of each, in distinct LDS arrays
// selected at compile time, so that reads of one buffer never wait for copies into the other);
// the rounds of a step then read only LDS.  Each step begins with one wait for the copies and
// one barrier.  A target's hops are added in increasing source order (steps in order, rows in
// order), the order of the CSR lists, so every H*V kernel gives the same bits.
// The down pass writes y = part(down hops); the up pass adds part(up hops) and part(diag x) to it
// (in the order of hv_finish) and takes the <xs x | y> partials of the finished y.

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(4))) const uint32_t const_u32_t;

// Read-only 16-byte record at p[i] through the constant address space: with a wave-uniform
// index the load is a scalar-memory load (counted apart from vector-memory loads).
__device__ __forceinline__ uint4 load_const4(const uint4* p, int64_t i) {
  const const_u32_t* q = (const const_u32_t*)(p + i);
  return make_uint4(q[0], q[1], q[2], q[3]);
}

// Buffer resource for raw byte offsets from a wave-uniform base (stride 0, every 32-bit offset
// in range; data format 32 bits, the word that the gfx9 family expects).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), static_cast<short>(0), 0x7fffffff,
                                           0x00020000);
}

// Asynchronous copy of 16 bytes from base + voff to LDS: the wavefront's active lanes write
// lds_wave + lane * 16 (lds_wave must be the same for the whole wavefront).  The buffer form
// (not global_load_lds) matters for the waits: the compiler counts these copies in vmcnt only,
// while a global_load_lds in flight makes it wait for every LDS read with lgkmcnt(0).
__device__ __forceinline__ void lds_copy16(__amdgpu_buffer_rsrc_t rsrc, uint32_t voff, void* lds_wave) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void_t*)lds_wave, 16, voff, 0, 0, 0);
}

// The same for 4 bytes (lane l writes lds_wave + l * 4).
__device__ __forceinline__ void lds_copy4(__amdgpu_buffer_rsrc_t rsrc, uint32_t voff, void* lds_wave) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void_t*)lds_wave, 4, voff, 0, 0, 0);
}

// Workgroup barrier without a memory fence: __syncthreads() would also wait for every copy to
// LDS in flight (its release covers them), which the staged epilogue keeps in flight.  Callers
// order LDS themselves: each wavefront waits for its own copies (vmcnt) or reads (lgkmcnt)
// before it.  The empty asm statements keep the compiler from moving memory accesses across.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Waits until every outstanding memory operation of the wavefront has completed.
__device__ __forceinline__ void wait_all_memory() { __builtin_amdgcn_s_waitcnt(0); }

template <typename T, bool REAL, bool UP>
__global__ __launch_bounds__(kStepThreads) void k_hv_step(const SectorView S, const typename Cx<T>::v2* __restrict__ x,
                                                          typename Cx<T>::v2* __restrict__ y, const int RT, const int r0,
                                                          const double* __restrict__ xs, double2* __restrict__ dotp,
                                                          const int64_t slot0, int* __restrict__ ctr, const int nblk,
                                                          const int ntiles) {
  using V2 = typename Cx<T>::v2;
  constexpr int VPL = 32 / static_cast<int>(sizeof(V2));  // vectors per lane: a lane owns 32 bytes
  constexpr int RB = 4 * VPL;                             // vectors per launch (128-byte elements)
  constexpr int NR = kStepRounds;
  constexpr int kEw = kStepRows * kStepGroup;  // entries of one wavefront's buffer
  __shared__ uint4 s_x0[kStepSub * 8];
  __shared__ uint4 s_x1[kStepSub * 8];
  __shared__ uint32_t s_e0[kStepWaves * kEw];
  __shared__ uint32_t s_e1[kStepWaves * kEw];
  __shared__ uint4 s_c[kStepCoefSlots];  // line-signed coefficients [parity * 256 + bond], sizeof(V2) bytes apart
  __shared__ int s_tile;
  __shared__ double s_xs[RB];  // xs of the launch's vectors (1 without xs)
  const int D = UP ? S.Dup : S.Ddn;
  const int32_t* __restrict__ blkp = UP ? S.st_blk_up : S.st_blk_dn;
  // step descriptors are scalar loads (load_const4), so that taking one never waits for a copy
  const uint4* __restrict__ desc = UP ? S.st_desc_up : S.st_desc_dn;
  const char* __restrict__ sent = reinterpret_cast<const char*>(UP ? S.st_ent_up : S.st_ent_dn);
  const double2* __restrict__ coef = UP ? S.coef_up : S.coef_dn;
  const uint32_t* __restrict__ xmask = UP ? S.xmask_up : S.xmask_dn;
  const uint32_t* __restrict__ lmask = UP ? S.dn_mask : S.up_mask;  // other spin's string of a line
  const int nbond = UP ? S.nbond_up : S.nbond_dn;
  const int64_t lstride = UP ? static_cast<int64_t>(RT) : static_cast<int64_t>(S.Ddn) * RT;
  const int64_t pstride = UP ? static_cast<int64_t>(S.Ddn) * RT : static_cast<int64_t>(RT);
  const int64_t pbytes = pstride * static_cast<int64_t>(sizeof(V2));
  const int xcd = blockIdx.x % kXcd;
  const int tid = static_cast<int>(threadIdx.x);
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int sub = tid & 3, el = tid >> 2, gl = el & (kStepGroup - 1);
  // A lane owns the two 16-byte pieces 2 sub and 2 sub + 1 of its target's elements and reads
  // them with two 16-byte LDS reads.  Each 8-lane group of such a read serves two targets; the
  // odd target takes its pieces in the other order, so the group covers all 32 banks whatever
  // the two source elements.  Accumulator slot v of an odd target holds vector v ^ H.
  constexpr int H = VPL / 2;
  const int odd = gl & 1, hx = odd * H;
  const int per = (ntiles + kXcd - 1) / kXcd;
  const int t0 = min(ntiles, xcd * per), nmine = min(ntiles, t0 + per) - t0;
  double run_re = 0.0, run_im = 0.0;  // threads < RB: this workgroup's <xs x | y> of vector r0 + tid
  if (tid < RB) s_xs[tid] = xs ? xs[r0 + tid] : 1.0;  // ordered before use by the first tile's barriers
  int cur_line = -1;
  for (int local = next_tile(ctr, xcd, &s_tile); local < nmine; local = next_tile(ctr, xcd, &s_tile)) {
    const int tile = __builtin_amdgcn_readfirstlane(t0 + local);
    const int line = tile / nblk;
    const int blk = tile - line * nblk;
    if (line != cur_line) {  // uniform; the first step's barrier orders these writes
      const uint32_t om = !UP && S.line_umask ? S.line_umask[line] : lmask[line];
      for (int b = tid; b < kStepCoefSlots / 2; b += kStepThreads) {
        V2 c;
        c.x = c.y = static_cast<T>(0);
        if (b < nbond) {
          const double2 cc = coef[b];
          const bool neg = (__popc(om & xmask[b]) & 1u) != 0;
          c.x = static_cast<T>(neg ? -cc.x : cc.x);
          c.y = static_cast<T>(neg ? -cc.y : cc.y);
        }
        V2* sc = reinterpret_cast<V2*>(s_c);  // slots sizeof(V2) bytes apart
        sc[b] = c;
        c.x = -c.x;
        c.y = -c.y;
        sc[b + kStepCoefSlots / 2] = c;
      }
      cur_line = line;
    }
    // element p of the line at xl + p * pstride
    const V2* __restrict__ xl = x + (!UP && S.line_base ? S.line_base[line] * RT : line * lstride) + r0;
    const int te = blk * kStepBlock + el;                 // this thread's target of round 0
    T ar[NR][VPL], ai[NR][VPL];
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
      for (int v = 0; v < VPL; ++v) ar[j][v] = ai[j][v] = static_cast<T>(0);
    const int st0 = __builtin_amdgcn_readfirstlane(blkp[blk]);
    const int st1 = __builtin_amdgcn_readfirstlane(blkp[blk + 1]);
    // copies of one step: thread tid moves 16-byte piece tid % 8 of source element tid / 8 (plus
    // kStepThreads / 8 elements per further kStepThreads pieces); lane l of wavefront w moves
    // 16 bytes l (and 64 + l) of w's rows.  Offsets from the step's first source element stay
    // below kStepSub * pbytes, under 2^31 (checked by the launcher).
    const uint32_t xlane = static_cast<uint32_t>((tid >> 3) * pbytes + (tid & 7) * 16);
    auto issue = [&](const uint4 d, uint4* xb, uint32_t* eb) {
      const uint32_t xoff = d.x & 0xfffffu, nu = (d.x >> 20) * 8u;
      const uint32_t row0 = d.y & 0x3ffffffu, ne = (d.y >> 26) * 4u;
      const __amdgpu_buffer_rsrc_t rx = raw_rsrc(reinterpret_cast<const char*>(xl) + static_cast<int64_t>(xoff) * pbytes);
#pragma unroll
      for (int it = 0; it < (kStepSub * 8 + kStepThreads - 1) / kStepThreads; ++it)
        if (static_cast<uint32_t>(it * kStepThreads + tid) < nu)
          lds_copy16(rx, xlane + static_cast<uint32_t>(it * (kStepThreads / 8) * pbytes),
                     reinterpret_cast<char*>(xb) + (it * kStepThreads + w * 64) * 16);
      const __amdgpu_buffer_rsrc_t re = raw_rsrc(sent + static_cast<int64_t>(row0) * (kStepGroup * 4));
      char* ed = reinterpret_cast<char*>(eb + w * kEw);
      if (static_cast<uint32_t>(lane) < ne) lds_copy16(re, lane * 16u, ed);
      if (static_cast<uint32_t>(lane + 64) < ne) lds_copy16(re, lane * 16u + 1024u, ed + 1024);
    };
    // one hop: its source's two 16-byte pieces and its coefficient (fetch), then the FMAs
    struct HopData {
      V2 x[VPL];
      V2 c;
    };
    auto fetch = [&](HopData& h, const char* xb1, const char* xb2, uint32_t en) {
      // entry = source byte offset | coefficient slot * 8 << 16 (16-bit halves)
      const uint32_t eo = en & 0xffffu;
      // each piece is one 16-byte LDS read (uint4 keeps the compiler from splitting it)
      const uint4 q1 = *reinterpret_cast<const uint4*>(xb1 + eo);
      const uint4 q2 = *reinterpret_cast<const uint4*>(xb2 + eo);
      static_assert(sizeof(V2) * H == sizeof(uint4), "a piece is H vectors");
      __builtin_memcpy(&h.x[0], &q1, sizeof(uint4));
      __builtin_memcpy(&h.x[H], &q2, sizeof(uint4));
      // the entry holds the coefficient slot times 8: a byte offset for complex64 slots (8 bytes
      // apart, twice the bank positions of 16-byte slots), half of one for complex128
      const uint32_t co = sizeof(V2) == 8 ? (en >> 16) : (en >> 16) << 1;
      h.c = *static_cast<const V2*>(
          __builtin_assume_aligned(reinterpret_cast<const char*>(s_c) + co, sizeof(V2)));
    };
    auto fmas = [&](T (&cr)[VPL], T (&cim)[VPL], const HopData& h) {
#pragma unroll
      for (int v = 0; v < VPL; ++v) {
        const V2 x2 = h.x[v];
        cr[v] = fmaT(h.c.x, x2.x, cr[v]);
        cim[v] = fmaT(h.c.x, x2.y, cim[v]);
        if (!REAL) {
          cr[v] = fmaT(-h.c.y, x2.y, cr[v]);
          cim[v] = fmaT(h.c.y, x2.x, cim[v]);
        }
      }
    };
    // Column pass: the epilogue's inputs of each round (y of the row pass and x at the round's
    // 256 targets, and the diagonal's up-string tables) are staged in LDS two rounds ahead, in
    // the step buffers (free once the last step is done): round j in set j % 2, y in s_x (32 KB)
    // with the tables behind it (planes 4 KB, up strings 1 KB, nuu 1 KB), x in s_e (32 KB).  A
    // target past D copies element D - 1 instead (its lanes store nothing), so every wavefront
    // issues the same copies: 4 of y and x, and one of a table in wavefronts 0-11.
    static_assert(sizeof(s_x0) >= kStepRound * (128 + 24) && sizeof(s_e0) >= kStepRound * 128,
                  "a round's staged inputs fit the step buffers");
    const int rt0 = blk * kStepBlock;
    const int nround = min(NR, (D - rt0 + kStepRound - 1) / kStepRound);  // rounds with a target
    const V2* __restrict__ yl = y + (xl - x);
    const uint32_t pb32 = static_cast<uint32_t>(pbytes);
    auto stage = [&](const int j, uint4* sxs, uint32_t* ses) {
      int tq = static_cast<int>(threadIdx.x);  // per-lane values are recomputed, not kept live
      asm volatile("" : "+v"(tq));
      const int rt = rt0 + j * kStepRound;  // below D for every staged round
      const uint32_t kmax = static_cast<uint32_t>(D - 1 - rt);
      const int64_t eo0 = static_cast<int64_t>(rt) * pstride;
      const __amdgpu_buffer_rsrc_t ry = raw_rsrc(yl + eo0), rx = raw_rsrc(xl + eo0);
#pragma unroll
      for (int it = 0; it < kStepRound * 8 / kStepThreads; ++it) {  // y and x: 4 copies per thread
        const uint32_t k = min(static_cast<uint32_t>(it * (kStepThreads / 8) + (tq >> 3)), kmax);
        const uint32_t vo = k * pb32 + static_cast<uint32_t>(tq & 7) * 16u;
        const int lo = (it * kStepThreads + w * 64) * 16;
        lds_copy16(ry, vo, reinterpret_cast<char*>(sxs) + lo);
        lds_copy16(rx, vo, reinterpret_cast<char*>(ses) + lo);
      }
      const uint32_t kd = min(static_cast<uint32_t>(tq & (kStepRound - 1)), kmax);
      char* sd = reinterpret_cast<char*>(sxs) + kStepRound * 128;
      if (w < 4)
        lds_copy16(raw_rsrc(S.planes + rt), kd * 16u, sd + w * 1024);
      else if (w < 8)
        lds_copy4(raw_rsrc(S.up_mask + rt), kd * 4u, sd + 4096 + (w - 4) * 256);
      else if (w < 12)
        lds_copy4(raw_rsrc(S.nuu + rt), kd * 4u, sd + 5120 + (w - 8) * 256);
    };
    uint4 dcur = make_uint4(0u, 0u, 0u, 0u), dnext = dcur;
    if (st0 < st1) {
      dcur = load_const4(desc, static_cast<int64_t>(st0) * kStepWaves + w);
      issue(dcur, s_x0, s_e0);
    }
    if (st0 + 1 < st1) dnext = load_const4(desc, static_cast<int64_t>(st0 + 1) * kStepWaves + w);
    // step st with buffers P (compile-time): wait for its copies, start the next step's copies,
    // then the rounds
    auto step = [&](auto pc, const int st) {
      constexpr int P = decltype(pc)::value;
      // the descriptor of step st + 2 is loaded before the wait, so that no scalar load is in
      // flight during the rounds: LDS reads alone then complete in order and each hop waits only
      // for its own reads
      uint4 dn2 = dnext;
      if (st + 2 < st1) dn2 = load_const4(desc, static_cast<int64_t>(st + 2) * kStepWaves + w);
      wait_all_memory();
      __syncthreads();  // step st has landed; every wavefront is done with the other buffers
      const uint4 d = dcur;
      if (st + 1 < st1) issue(dnext, P ? s_x0 : s_x1, P ? s_e0 : s_e1);
      dcur = dnext;
      dnext = dn2;
      const char* xb1 = reinterpret_cast<const char*>(P ? s_x1 : s_x0) + (sub * 32 + odd * 16);
      const char* xb2 = reinterpret_cast<const char*>(P ? s_x1 : s_x0) + (sub * 32 + (16 - odd * 16));
      // one offset for the whole per-lane base: the compiler's LDS alias information follows only
      // a few levels of address arithmetic, and reads it loses wait for every pending copy
      const uint32_t* eb = (P ? s_e1 : s_e0) + (w * kEw + gl);
      const uint32_t c_lo = __builtin_amdgcn_readfirstlane(d.z), c_hi = __builtin_amdgcn_readfirstlane(d.w);
      // row r of this wavefront's entries at eb[r * 16]; the row offsets pass through
      // readfirstlane so that each read stays one offset away from eb (see above).  Rows are
      // read two ahead of their hops (e0, e1 = rows r, r + 1); reads past the last row are
      // clamped to row kStepRows - 1 and never used.
      auto row = [&](int r) { return eb[__builtin_amdgcn_readfirstlane(min(r, kStepRows - 1) * kStepGroup)]; };
      uint32_t e0 = row(0), e1 = row(1);
      int r = 0;
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const int cnt = static_cast<int>(((j < 4 ? c_lo : c_hi) >> (8 * (j & 3))) & 0xffu);
        for (int k = 0; k + 2 <= cnt; k += 2) {
          HopData ha, hb;
          fetch(ha, xb1, xb2, e0);
          fetch(hb, xb1, xb2, e1);
          e0 = row(r + 2);
          e1 = row(r + 3);
          fmas(ar[j], ai[j], ha);
          fmas(ar[j], ai[j], hb);
          r += 2;
        }
        if (cnt & 1) {
          HopData ha;
          fetch(ha, xb1, xb2, e0);
          e0 = e1;
          e1 = row(r + 2);
          fmas(ar[j], ai[j], ha);
          r += 1;
        }
      }
    };
    for (int st = st0; st < st1; st += 2) {
      step(std::integral_constant<int, 0>{}, st);
      if (st + 1 < st1) step(std::integral_constant<int, 1>{}, st + 1);
    }
    // epilogue.  Down: y = part(S_down).  Up: y = (y + part(S_up)) + part(diag * x), and the
    // <xs x | y> partials of the finished y.  The opaque copy of the lane's offset keeps the
    // compiler from holding addresses live across the steps.
    int64_t eoff = static_cast<int64_t>(te) * pstride + sub * VPL;
    asm volatile("" : "+v"(eoff));
    const V2* xta = xl + eoff + hx;        // slots [0, H)
    const V2* xtb = xl + eoff + (H - hx);  // slots [H, VPL)
    V2* yta = y + (xta - x);
    V2* ytb = y + (xtb - x);
    // xs of slot v (vector v ^ hx of the lane's piece), read from LDS where used
    const double* sxl = s_xs + sub * VPL;
    double dre[VPL], dimg[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) dre[v] = dimg[v] = 0.0;
    if constexpr (UP) {
      const uint32_t dmask = lmask[line];  // the line's down string
      const int nd_line = S.ndd[line];
      wait_all_memory();
      __syncthreads();  // every wavefront is done with both step buffers
      stage(0, s_x0, s_e0);
      if (nround > 1) stage(1, s_x1, s_e1);
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        if (j >= nround) continue;  // uniform
        // round j's copies have landed in every wavefront: this wavefront's own copies are older
        // than the next round's 4 or 5 (when there is a next round), which may stay in flight
        if (j + 1 < nround)
          __builtin_amdgcn_s_waitcnt(0x0f74);  // vmcnt(4)
        else
          __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
        lds_barrier();
        // per-lane values from an opaque copy of the thread index, so that none stays live
        // across rounds (registers are short here, and a spill reload waits for the copies)
        int tq = static_cast<int>(threadIdx.x);
        asm volatile("" : "+v"(tq));
        const int eq = tq >> 2, pq = (tq & 3) * 32 + ((eq & 1) ? 16 : 0), pq2 = (tq & 3) * 32 + ((eq & 1) ? 0 : 16);
        const bool q1 = (j & 1) != 0;  // compile-time: the loop is unrolled
        const char* sy = reinterpret_cast<const char*>(q1 ? s_x1 : s_x0) + eq * 128;
        const char* sxe = reinterpret_cast<const char*>(q1 ? s_e1 : s_e0) + eq * 128;
        const char* sd = reinterpret_cast<const char*>(q1 ? s_x1 : s_x0) + kStepRound * 128;
        const int t = rt0 + j * kStepRound + eq;
        const bool live = t < D;
        V2 py[VPL], px[VPL];
        uint4 pl = make_uint4(0u, 0u, 0u, 0u);
        uint32_t um = 0u;
        int nu = 0;
        if (live) {
          const uint4 y1 = *reinterpret_cast<const uint4*>(sy + pq);
          const uint4 y2 = *reinterpret_cast<const uint4*>(sy + pq2);
          const uint4 x1 = *reinterpret_cast<const uint4*>(sxe + pq);
          const uint4 x2 = *reinterpret_cast<const uint4*>(sxe + pq2);
          __builtin_memcpy(&py[0], &y1, sizeof(uint4));
          __builtin_memcpy(&py[H], &y2, sizeof(uint4));
          __builtin_memcpy(&px[0], &x1, sizeof(uint4));
          __builtin_memcpy(&px[H], &x2, sizeof(uint4));
          pl = *reinterpret_cast<const uint4*>(sd + eq * 16);
          um = *reinterpret_cast<const uint32_t*>(sd + 4096 + eq * 4);
          nu = *reinterpret_cast<const int*>(sd + 5120 + eq * 4);
        }
        if (live) {
          const T dgt = static_cast<T>(diag_from(S, um, dmask, pl, nu, nd_line));
          // the element's two pieces in y (slots [0, H) at byte pq, [H, VPL) at byte pq2)
          char* ye = reinterpret_cast<char*>(const_cast<V2*>(yl + static_cast<int64_t>(t) * pstride));
          V2* ya = reinterpret_cast<V2*>(ye + pq);
          V2* yb = reinterpret_cast<V2*>(ye + pq2);
          const double* sxq = s_xs + (tq & 3) * VPL;
          const int hq = (eq & 1) ? H : 0;
#pragma unroll
          for (int v = 0; v < VPL; ++v) {
            V2* yp = v < H ? ya + v : yb + (v - H);
            const double sx = sxq[v ^ hq];
            T a_r = ar[j][v], a_i = ai[j][v];
            asm volatile("" : "+v"(a_r), "+v"(a_i));  // convert round by round (not all rounds up front)
            const V2 q = px[v];
            V2 o;
            o.x = hv_finish(py[v].x, a_r, dgt, q.x, sx);
            o.y = hv_finish(py[v].y, a_i, dgt, q.y, sx);
            *yp = o;
            if (dotp) {
              const double qr = sx * static_cast<double>(q.x), qi = sx * static_cast<double>(q.y);
              dre[v] += qr * static_cast<double>(o.x) + qi * static_cast<double>(o.y);
              dimg[v] += qr * static_cast<double>(o.y) - qi * static_cast<double>(o.x);
            }
          }
        }
        if (j + 2 < nround) {  // round j + 2 into the set every wavefront has finished reading
          __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wavefront's reads of round j are done
          lds_barrier();                        // and so are every wavefront's
          if (q1)
            stage(j + 2, s_x1, s_e1);
          else
            stage(j + 2, s_x0, s_e0);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        if (te + j * kStepRound >= D) continue;
        const int64_t jo = static_cast<int64_t>(j * kStepRound) * pstride;
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
          V2* yp = v < H ? yta + v : ytb + (v - H);
          T a_r = ar[j][v], a_i = ai[j][v];
          asm volatile("" : "+v"(a_r), "+v"(a_i));
          V2 o;
          o.x = hv_part(a_r, sxl[v ^ hx]);
          o.y = hv_part(a_i, sxl[v ^ hx]);
          yp[jo] = o;
        }
        asm volatile("" ::: "memory");
      }
    }
    if (dotp) {
      // slots back to vector order, then lanes of one wavefront with the same sub hold partials
      // of the same vectors: fold them (offsets 4..32) and reduce the 16 wavefronts' sums in LDS
#pragma unroll
      for (int v = 0; v < H; ++v) {
        const double a = dre[v], b = dimg[v], c2 = dre[v + H], d2 = dimg[v + H];
        dre[v] = odd ? c2 : a;
        dimg[v] = odd ? d2 : b;
        dre[v + H] = odd ? a : c2;
        dimg[v + H] = odd ? b : d2;
      }
#pragma unroll
      for (int v = 0; v < VPL; ++v)
#pragma unroll
        for (int m = 4; m < 64; m <<= 1) {
          dre[v] += __shfl_xor(dre[v], m);
          dimg[v] += __shfl_xor(dimg[v], m);
        }
      double* red = reinterpret_cast<double*>(s_e0);  // [16 waves][4 subs][VPL][2]
      __syncthreads();  // every wavefront is done with the last step's entries
      if (lane < 4) {
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
          red[((w * 4 + lane) * VPL + v) * 2] = dre[v];
          red[((w * 4 + lane) * VPL + v) * 2 + 1] = dimg[v];
        }
      }
      __syncthreads();
      if (tid < RB) {
        const int s2 = tid / VPL, v2 = tid % VPL;
        for (int ww = 0; ww < kStepWaves; ++ww) {
          run_re += red[((ww * 4 + s2) * VPL + v2) * 2];
          run_im += red[((ww * 4 + s2) * VPL + v2) * 2 + 1];
        }
      }
    }
  }
  queue_done(ctr);
  if (dotp && tid < RB) dotp[(slot0 + blockIdx.x) * RT + r0 + tid] = make_double2(run_re, run_im);
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
    const int64_t rowbase = static_cast<int64_t>(iu) * S.Ddn;
    for (int k = S.off_dn[id]; k < S.off_dn[id + 1]; ++k) {
      const uint32_t ent = S.ent_dn[k];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const uint32_t par = ((ent >> kParityBit) & 1u) ^ (__popc(u & S.xmask_dn[b]) & 1u);
      const double2 c = S.coef_dn[b];
      axpy_block<T, RB, REAL>(ar, ai, x + (rowbase + t) * RT + r0, static_cast<T>(par ? -c.x : c.x),
                              static_cast<T>(par ? -c.y : c.y));
    }
    V2* ye = y + e * RT + r0;
    T ur[RB], ui[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const double sx = xs ? xs[r0 + r] : 1.0;
      ur[r] = static_cast<T>(static_cast<double>(ar[r]) * sx);
      ui[r] = static_cast<T>(static_cast<double>(ai[r]) * sx);
      ar[r] = ai[r] = static_cast<T>(0);
    }
    // up bonds accumulate separately and are added last, as the column sweep does
    for (int k = S.off_up[iu]; k < S.off_up[iu + 1]; ++k) {
      const uint32_t ent = S.ent_up[k];
      const uint32_t t = ent & kTargetMask;
      const uint32_t b = (ent >> kTargetBits) & kBondMask;
      const uint32_t par = ((ent >> kParityBit) & 1u) ^ (__popc(d & S.xmask_up[b]) & 1u);
      const double2 c = S.coef_up[b];
      axpy_block<T, RB, REAL>(ar, ai, x + (static_cast<int64_t>(t) * S.Ddn + id) * RT + r0,
                              static_cast<T>(par ? -c.x : c.x), static_cast<T>(par ? -c.y : c.y));
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const double sx = xs ? xs[r0 + r] : 1.0;
      V2 o;
      o.x = ur[r] + static_cast<T>(static_cast<double>(ar[r]) * sx);
      o.y = ui[r] + static_cast<T>(static_cast<double>(ai[r]) * sx);
      ye[r] = o;
      if (dotp) {
        const double qr = sx * static_cast<double>(xo[r].x), qi = sx * static_cast<double>(xo[r].y);
        dre[r] += qr * static_cast<double>(o.x) + qi * static_cast<double>(o.y);
        dimg[r] += qr * static_cast<double>(o.y) - qi * static_cast<double>(o.x);
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

// dst[0 .. n) = 0 (a kernel rather than a memset node, so that applies can be graph-captured)
__global__ void k_zero_d2(double2* __restrict__ dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = make_double2(0.0, 0.0);
}

// ------------------------------------------------------------------ host dispatch
namespace {

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : dflt;
}

// Per-pass timing (ftlm_hip_apply_pass_times): while set, launch() records an event before the
// row pass, between the passes and after the column pass (a flat launch counts as a row pass).
thread_local std::vector<std::array<hipEvent_t, 3>>* t_pass_events = nullptr;

void record_pass_event(int slot, hipStream_t st) {
  if (!t_pass_events) return;
  if (slot == 0) t_pass_events->push_back({nullptr, nullptr, nullptr});
  hipEvent_t& e = t_pass_events->back()[slot];
  FTLM_HIP_CHECK(hipEventCreate(&e));
  FTLM_HIP_CHECK(hipEventRecord(e, st));
}

struct Plan {
  bool two = false;
  int grid_flat = 0;
  int grid_dn = 0, nchunk = 1, ntiles_dn = 0;
  int grid_up = 0, wlog = 0, tps = 1, ntiles_up = 0;
};

Plan make_plan(const Sector* Sx, int RT, int esize, int L, const HvShape& sh) {
  const SectorView S = Sx->view();
  Plan p;
  p.two = S.Ddn >= 128;
  if (!p.two) {
    const int64_t tiles = (S.dim + kHvBlock - 1) / kHvBlock;
    p.grid_flat = static_cast<int>(std::min<int64_t>(tiles, sh.grid > 0 ? std::min(sh.grid, kHvMaxGrid) : kHvMaxGrid));
    return p;
  }
  const int EPB = kHvBlock / L;
  const int64_t budget = static_cast<int64_t>(env_int("FTLM_HV_L2_KB", 3072)) * 1024;
  // workgroups per XCD: enough to fill its 32 CUs several times over
  const int bpx =
      std::max(1, std::min(kPassMaxGrid / kXcd, sh.bpx > 0 ? sh.bpx : env_int("FTLM_HV_BPX", 256)));
  p.nchunk = (S.Ddn + EPB - 1) / EPB;
  const int64_t ntd = static_cast<int64_t>(S.Dup) * p.nchunk;
  FTLM_REQUIRE(ntd < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "too many row tiles");
  p.ntiles_dn = static_cast<int>(ntd);
  p.grid_dn = kXcd * static_cast<int>(std::min<int64_t>(bpx, (ntd + kXcd - 1) / kXcd));
  // column pass: widest power-of-two stripe W <= EPB with D_up * W * RT * esize <= budget
  const int64_t col_bytes = static_cast<int64_t>(S.Dup) * RT * esize;
  int wl = 0;
  while ((1 << (wl + 1)) <= EPB && col_bytes * (int64_t{1} << (wl + 1)) <= budget) ++wl;
  wl = std::max(0, std::min(wl, env_int("FTLM_HV_WLOG", wl)));
  if (sh.wlog >= 0) wl = sh.wlog;
  while ((1 << wl) > EPB) --wl;
  p.wlog = wl;
  const int rows_per_tile = EPB >> wl;
  p.tps = (S.Dup + rows_per_tile - 1) / rows_per_tile;
  const int64_t ntu = static_cast<int64_t>((S.Ddn + (1 << wl) - 1) >> wl) * p.tps;
  FTLM_REQUIRE(ntu < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "too many column tiles");
  p.ntiles_up = static_cast<int>(ntu);
  p.grid_up = kXcd * static_cast<int>(std::min<int64_t>(bpx, (ntu + kXcd - 1) / kXcd));
  return p;
}

// Whether a two-pass sector apply of 128-byte elements uses the LDS-staged step passes (default)
// or the L2-gather passes k_hv_dn / k_hv_up (FTLM_HV_STEP=0, kept for comparison).
// The copies address a step's sources by 32-bit offsets from its first source element: kStepSub
// elements of the column pass (stride D_dn * RT vectors) must span less than 2^31 bytes.
bool use_step(const SectorView& S, int RT, size_t vbytes) {
  static const bool on = env_int("FTLM_HV_STEP", 1) != 0;
  const bool fits = static_cast<int64_t>(kStepSub) * S.Ddn * RT * static_cast<int64_t>(vbytes) < (int64_t{1} << 31);
  return on && fits && S.st_ent_up != nullptr && S.st_ent_dn != nullptr;
}

template <typename T, bool REAL, bool UP>
void launch_step_pass(const SectorView& S, const typename Cx<T>::v2* x, typename Cx<T>::v2* y, int RT, int r0,
                      const double* xs, double2* dotp, int64_t slot0, int* ctr, int grid, hipStream_t st) {
  const int D = UP ? S.Dup : S.Ddn, nline = UP ? S.Ddn : S.Dup;
  const int nblk = (D + kStepBlock - 1) / kStepBlock;
  const int64_t nt = static_cast<int64_t>(nline) * nblk;
  FTLM_REQUIRE(nt < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "too many step tiles");
  hipLaunchKernelGGL((k_hv_step<T, REAL, UP>), dim3(grid), dim3(kStepThreads), 0, st, S, x, y, RT, r0, xs, dotp,
                     slot0, ctr, nblk, static_cast<int>(nt));
}

template <typename T, int RB, bool REAL>
int launch(const Sector* Sx, const void* x, void* y, int RT, int r0, const double* xs, double2* dotp,
           hipStream_t st, const HvShape& sh) {
  using V2 = typename Cx<T>::v2;
  using SP = Split<T, RB>;
  const SectorView S = Sx->view();
  const Plan p = make_plan(Sx, RT, sizeof(V2), SP::L, sh);
  const V2* xp = static_cast<const V2*>(x);
  V2* yp = static_cast<V2*>(y);
  if constexpr (RB * sizeof(V2) == 128) {  // 16 complex64 or 8 complex128 vectors
    if (p.two && use_step(S, RT, sizeof(V2))) {
      // as many workgroups per CU as the step buffers fit in its LDS (32 CUs per XCD)
      const int bpx = std::max(1, std::min(32 * kStepWgPerCu, sh.bpx > 0 ? sh.bpx : 32 * kStepWgPerCu));
      const int grid = kXcd * bpx;
      int32_t* ctr = Sx->tile_counters(st);
      // <xs x | y> comes whole from the column pass, which reads x at the targets anyway: the
      // row pass would have to read x there a second time for its part
      record_pass_event(0, st);
      launch_step_pass<T, REAL, false>(S, xp, yp, RT, r0, xs, nullptr, 0, ctr, grid, st);
      record_pass_event(1, st);
      launch_step_pass<T, REAL, true>(S, xp, yp, RT, r0, xs, dotp, 0, ctr + 16, grid, st);
      record_pass_event(2, st);
      return grid;
    }
  }
  if (!p.two) {
    record_pass_event(0, st);
    hipLaunchKernelGGL((k_hv_flat<T, RB, REAL>), dim3(p.grid_flat), dim3(kHvBlock), 0, st, S, xp, yp, RT, r0, xs,
                       dotp);
    record_pass_event(1, st);
    record_pass_event(2, st);
    return p.grid_flat;
  }
  const bool stage = StageEntries<SP::EPB>::value;
  const size_t lds_dn = stage ? static_cast<size_t>(std::max(S.maxcnt_dn, 1)) * SP::EPB * sizeof(uint32_t) : 16;
  const size_t lds_up =
      stage ? static_cast<size_t>(std::max(S.maxcnt_up, 1)) * (SP::EPB >> p.wlog) * sizeof(uint32_t) : 16;
  FTLM_REQUIRE(lds_dn <= 48 * 1024 && lds_up <= 48 * 1024, FTLM_EUNSUPPORTED,
               "hop table of one H*V tile exceeds 48 KB of LDS");
  FTLM_REQUIRE(static_cast<int64_t>(S.Ddn) * RT < (int64_t{1} << 24), FTLM_EUNSUPPORTED,
               "D_dn * R must stay below 2^24 for the column pass's 24-bit source offsets");
  int32_t* ctr = Sx->tile_counters(st);
  record_pass_event(0, st);
  // the row pass writes part(S_down); the column pass adds the rest and takes the <xs x | y>
  // partials of the finished y, reading x at its targets for the diagonal anyway
  hipLaunchKernelGGL((k_hv_dn<T, RB, REAL>), dim3(p.grid_dn), dim3(kHvBlock), lds_dn, st, S, xp, yp, RT, r0, xs,
                     ctr, p.nchunk, p.ntiles_dn);
  record_pass_event(1, st);
  hipLaunchKernelGGL((k_hv_up<T, RB, REAL>), dim3(p.grid_up), dim3(kHvBlock), lds_up, st, S, xp, yp, RT, r0, xs,
                     dotp, int64_t{0}, ctr + 16, p.wlog, p.tps, p.ntiles_up);
  record_pass_event(2, st);
  return p.grid_up;
}

template <typename T, bool REAL>
int launch_chunk(const Sector* S, const void* x, void* y, int RT, int r0, int rb, const double* xs, double2* dotp,
                 hipStream_t st, const HvShape& sh) {
  switch (rb) {
    case 16: return launch<T, 16, REAL>(S, x, y, RT, r0, xs, dotp, st, sh);
    case 8: return launch<T, 8, REAL>(S, x, y, RT, r0, xs, dotp, st, sh);
    case 4: return launch<T, 4, REAL>(S, x, y, RT, r0, xs, dotp, st, sh);
    case 2: return launch<T, 2, REAL>(S, x, y, RT, r0, xs, dotp, st, sh);
    default: return launch<T, 1, REAL>(S, x, y, RT, r0, xs, dotp, st, sh);
  }
}

}  // namespace

// Upper bound of the partial-sum blocks any H*V launch on this operator writes.
int hv_partial_blocks(const ftlm_op* op) {
  (void)op;
  return kHvMaxGrid;
}

int kblock_default_grid(const ftlm_op* op, int* grid_max);
int kstack_default_grid(const ftlm_op* op, int* grid_max);

int hv_default_grid(const ftlm_op* op, int* grid_max) {
  if (op->kind == kKBlock) return kblock_default_grid(op, grid_max);
  if (op->kind == kKStack) return kstack_default_grid(op, grid_max);
  if (grid_max) *grid_max = kHvMaxGrid;
  return static_cast<int>(std::min<int64_t>((op->dim + kHvBlock - 1) / kHvBlock, kHvMaxGrid));
}

int hv_default_wlog(const ftlm_op* op, int R, int dtype, int rb, int* wlog_max) {
  if (wlog_max) *wlog_max = -1;
  if (op->kind != kSector) return -1;
  const Sector* Sx = static_cast<const Sector*>(op);
  if (Sx->Ddn < 128) return -1;
  const int esize = dtype == FTLM_C64 ? static_cast<int>(sizeof(float2)) : static_cast<int>(sizeof(double2));
  const int bytes = rb * esize, L = bytes >= 32 ? bytes / 16 : 1;
  const Plan p = make_plan(Sx, R, esize, L, HvShape{});
  int wm = 0;
  while ((1 << (wm + 1)) <= kHvBlock / L) ++wm;
  if (wlog_max) *wlog_max = wm;
  return p.wlog;
}

bool sector_has_step_tables(const Sector* S) { return S->st_ent[0].p != nullptr && S->st_ent[1].p != nullptr; }

namespace {

// One chunk of RB vectors of hv_rows_lines: the step row pass for 128-byte elements, the L2-gather
// row pass otherwise (both add a target's down hops in the same order).
template <typename T, int RB>
void rows_lines_chunk(const Sector* Sx, const SectorView& S, const void* x, void* y, int RT, int r0, const double* xs,
                      int nlines, hipStream_t st) {
  using V2 = typename Cx<T>::v2;
  using SP = Split<T, RB>;
  int32_t* ctr = Sx->tile_counters(st);
  const V2* xp = static_cast<const V2*>(x);
  V2* yp = static_cast<V2*>(y);
  if constexpr (RB * sizeof(V2) == 128) {
    if (use_step(S, RT, sizeof(V2))) {
      const int nblk = (S.Ddn + kStepBlock - 1) / kStepBlock;
      const int64_t nt = static_cast<int64_t>(nlines) * nblk;
      FTLM_REQUIRE(nt < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "too many step tiles");
      hipLaunchKernelGGL((k_hv_step<T, true, false>), dim3(kXcd * 32 * kStepWgPerCu), dim3(kStepThreads), 0, st, S,
                         xp, yp, RT, r0, xs, nullptr, int64_t{0}, ctr, nblk, static_cast<int>(nt));
      return;
    }
  }
  const Plan p = make_plan(Sx, RT, sizeof(V2), SP::L, HvShape{});
  FTLM_REQUIRE(p.two, FTLM_EUNSUPPORTED, "row pass over lines needs D_dn >= 128");
  const bool stage = StageEntries<SP::EPB>::value;
  const size_t lds_dn = stage ? static_cast<size_t>(std::max(S.maxcnt_dn, 1)) * SP::EPB * sizeof(uint32_t) : 16;
  FTLM_REQUIRE(lds_dn <= 48 * 1024, FTLM_EUNSUPPORTED, "hop table of one H*V tile exceeds 48 KB of LDS");
  const int64_t ntd = static_cast<int64_t>(nlines) * p.nchunk;
  FTLM_REQUIRE(ntd < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "too many row tiles");
  const int grid = kXcd * static_cast<int>(std::min<int64_t>(p.grid_dn / kXcd, (ntd + kXcd - 1) / kXcd));
  hipLaunchKernelGGL((k_hv_dn<T, RB, true>), dim3(grid), dim3(kHvBlock), lds_dn, st, S, xp, yp, RT, r0, xs, ctr,
                     p.nchunk, static_cast<int>(ntd));
}

template <typename T>
void rows_lines_dispatch(const Sector* Sx, const SectorView& S, const void* x, void* y, int RT, int r0, int rb,
                         const double* xs, int nlines, hipStream_t st) {
  switch (rb) {
    case 16: rows_lines_chunk<T, 16>(Sx, S, x, y, RT, r0, xs, nlines, st); break;
    case 8: rows_lines_chunk<T, 8>(Sx, S, x, y, RT, r0, xs, nlines, st); break;
    case 4: rows_lines_chunk<T, 4>(Sx, S, x, y, RT, r0, xs, nlines, st); break;
    case 2: rows_lines_chunk<T, 2>(Sx, S, x, y, RT, r0, xs, nlines, st); break;
    default: rows_lines_chunk<T, 1>(Sx, S, x, y, RT, r0, xs, nlines, st); break;
  }
}

}  // namespace

void hv_rows_lines(const Sector* Sx, const void* x, void* y, int R, int dtype, const double* xs, const int64_t* lbase,
                   const uint32_t* lumask, int nlines, hipStream_t st) {
  FTLM_REQUIRE(Sx->real_h, FTLM_EUNSUPPORTED, "row pass over lines needs real bond coefficients");
  if (nlines <= 0) return;
  SectorView S = Sx->view();
  S.line_base = lbase;
  S.line_umask = lumask;
  const int maxrb = dtype == FTLM_C64 ? 16 : 8;
  for (int r0 = 0; r0 < R;) {
    int rb = 1;
    while (rb * 2 <= maxrb && r0 + rb * 2 <= R) rb *= 2;
    if (dtype == FTLM_C64)
      rows_lines_dispatch<float>(Sx, S, x, y, R, r0, rb, xs, nlines, st);
    else
      rows_lines_dispatch<double>(Sx, S, x, y, R, r0, rb, xs, nlines, st);
    r0 += rb;
  }
  check_launch("H*V row pass over lines");
}

// Block apply over all RT vectors in chunks of RB (<= 8 for complex128, <= 16 for complex64).
// Chunks are either all two-pass or all flat; the number of dot-partial blocks written is the
// maximum over chunks, and blocks a chunk does not launch are zero-filled first.
int hv_apply_sector(const Sector* S, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st, bool force_real) {
  const bool real = S->real_h || force_real;
  const HvShape sh = S->shape(R, dtype);
  const int maxrb = sh.rb > 0 ? std::min(sh.rb, dtype == FTLM_C64 ? 16 : 8) : (dtype == FTLM_C64 ? 16 : 8);
  int nblk = 0;
  std::vector<std::pair<int, int>> chunks;
  for (int r0 = 0; r0 < R;) {
    int rb = 1;
    while (rb * 2 <= maxrb && r0 + rb * 2 <= R) rb *= 2;
    chunks.emplace_back(r0, rb);
    r0 += rb;
  }
  if (dotp && chunks.size() > 1) {
    const int64_t nz = static_cast<int64_t>(kHvMaxGrid) * R;
    hipLaunchKernelGGL(k_zero_d2, dim3(static_cast<unsigned>(std::min<int64_t>(1024, (nz + 255) / 256))), dim3(256),
                       0, st, dotp, nz);
  }
  for (const auto& c : chunks) {
    int nb;
    if (dtype == FTLM_C64) {
      nb = real ? launch_chunk<float, true>(S, x, y, R, c.first, c.second, xs, dotp, st, sh)
                : launch_chunk<float, false>(S, x, y, R, c.first, c.second, xs, dotp, st, sh);
    } else {
      nb = real ? launch_chunk<double, true>(S, x, y, R, c.first, c.second, xs, dotp, st, sh)
                : launch_chunk<double, false>(S, x, y, R, c.first, c.second, xs, dotp, st, sh);
    }
    nblk = std::max(nblk, nb);
  }
  check_launch("k_hv_sector");
  return nblk;
}

int hv_apply_kblock(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st);
int hv_apply_callback(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs,
                      double2* dotp, hipStream_t st);
int hv_apply_gram(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                  hipStream_t st);
int hv_apply_kstack(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st);

int hv_apply(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
             hipStream_t st, bool force_real) {
  if (op->kind == kSector)
    return hv_apply_sector(static_cast<const Sector*>(op), x, y, R, dtype, xs, dotp, st, force_real);
  FTLM_REQUIRE(!force_real, FTLM_EUNSUPPORTED, "the real-part apply is defined for full-sector operators only");
  if (op->kind == kCallback) return hv_apply_callback(op, x, y, R, dtype, xs, dotp, st);
  if (op->kind == kGram) return hv_apply_gram(op, x, y, R, dtype, xs, dotp, st);
  if (op->kind == kKStack) return hv_apply_kstack(op, x, y, R, dtype, xs, dotp, st);
  return hv_apply_kblock(op, x, y, R, dtype, xs, dotp, st);
}

}  // namespace ftlm_hip

using namespace ftlm_hip;

namespace {

// memcpy of large host blocks on several threads (pageable <-> pinned staging runs at the speed
// of one core otherwise, below the PCIe rate).
void par_copy(void* dst, const void* src, size_t bytes) {
  constexpr size_t kPart = size_t{4} << 20;
  const int nt = static_cast<int>(std::min<size_t>(8, (bytes + kPart - 1) / kPart));
  if (nt <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  auto part = [&](int t) {
    const size_t a = bytes * t / nt, b = bytes * (t + 1) / nt;
    std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
  };
  for (int t = 1; t < nt; ++t) th.emplace_back(part, t);
  part(0);
  for (auto& x : th) x.join();
}

constexpr size_t kStageChunk = size_t{4} << 20;  // complex128 elements per pinned chunk (64 MB)

// y = H x for host vectors: x is copied into the operator's pinned chunks (on host threads) and
// on to the device while the next chunk is filled, the apply runs on the operator's own stream,
// and y comes back the same way.  The device blocks, pinned chunks, stream and events live with
// the operator (made on its first host apply).
void apply_host_staged(const ftlm_op* op, const double2* x, double2* y) {
  const size_t n = static_cast<size_t>(op->dim);
  if (n == 0) return;
  HostStage* S = nullptr;
  {
    std::lock_guard<std::mutex> lk(op->stage_mu);
    if (!op->host_stage) {
      auto st = std::make_unique<HostStage>();
      st->dx.alloc(n);
      st->dy.alloc(n);
      st->pin_n = std::min(n, kStageChunk);
      for (int i = 0; i < 2; ++i) {
        FTLM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&st->pin[i]), st->pin_n * sizeof(double2),
                                     hipHostMallocDefault));
        FTLM_HIP_CHECK(hipEventCreateWithFlags(&st->ev[i], hipEventDisableTiming));
      }
      FTLM_HIP_CHECK(hipStreamCreateWithFlags(&st->st, hipStreamNonBlocking));
      op->host_stage = std::move(st);
    }
    S = op->host_stage.get();
  }
  std::lock_guard<std::mutex> lk(S->mu);
  const size_t cn = S->pin_n, nc = (n + cn - 1) / cn;
  const size_t es = sizeof(double2);
  for (size_t c = 0; c < nc; ++c) {  // x: host -> pinned chunk c % 2 -> device
    const int b = static_cast<int>(c & 1);
    const size_t off = c * cn, len = std::min(cn, n - off);
    if (c >= 2) FTLM_HIP_CHECK(hipEventSynchronize(S->ev[b]));  // chunk c - 2's copy has left it
    par_copy(S->pin[b], x + off, len * es);
    FTLM_HIP_CHECK(hipMemcpyAsync(S->dx.p + off, S->pin[b], len * es, hipMemcpyHostToDevice, S->st));
    FTLM_HIP_CHECK(hipEventRecord(S->ev[b], S->st));
  }
  hv_apply(op, S->dx.p, S->dy.p, 1, FTLM_C128, nullptr, nullptr, S->st);
  auto fetch = [&](size_t c) {
    const int b = static_cast<int>(c & 1);
    const size_t off = c * cn, len = std::min(cn, n - off);
    FTLM_HIP_CHECK(hipMemcpyAsync(S->pin[b], S->dy.p + off, len * es, hipMemcpyDeviceToHost, S->st));
    FTLM_HIP_CHECK(hipEventRecord(S->ev[b], S->st));
  };
  fetch(0);
  for (size_t c = 0; c < nc; ++c) {  // y: device -> pinned chunk c % 2 -> host, one chunk ahead
    const int b = static_cast<int>(c & 1);
    const size_t off = c * cn, len = std::min(cn, n - off);
    FTLM_HIP_CHECK(hipEventSynchronize(S->ev[b]));
    if (c + 1 < nc) fetch(c + 1);  // the other chunk was emptied into y at step c - 1
    par_copy(y + off, S->pin[b], len * es);
  }
}

}  // namespace

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

int ftlm_hip_apply_pass_times(const ftlm_op_t* op, const void* x, void* y, int R, int dtype, void* stream,
                              double* ms_row, double* ms_col) {
  std::vector<std::array<hipEvent_t, 3>> ev;
  const int rc = guarded([&] {
    FTLM_REQUIRE(op != nullptr && x != nullptr && y != nullptr, FTLM_EINVAL, "null argument");
    FTLM_REQUIRE(R >= 1 && R <= 32, FTLM_EINVAL, "R must be in [1, 32]");
    FTLM_REQUIRE(dtype == FTLM_C128 || dtype == FTLM_C64, FTLM_EINVAL, "dtype must be FTLM_C128 or FTLM_C64");
    FTLM_REQUIRE(x != y, FTLM_EINVAL, "x and y must not alias");
    FTLM_REQUIRE(op->kind == kSector, FTLM_EUNSUPPORTED, "pass timing is defined for full-sector operators");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    t_pass_events = &ev;
    try {
      hv_apply(op, x, y, R, dtype, nullptr, nullptr, st);
    } catch (...) {
      t_pass_events = nullptr;
      throw;
    }
    t_pass_events = nullptr;
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
    double a = 0.0, b = 0.0;
    for (const auto& e : ev) {
      float t0 = 0.f, t1 = 0.f;
      FTLM_HIP_CHECK(hipEventElapsedTime(&t0, e[0], e[1]));
      FTLM_HIP_CHECK(hipEventElapsedTime(&t1, e[1], e[2]));
      a += t0;
      b += t1;
    }
    if (ms_row) *ms_row = a;
    if (ms_col) *ms_col = b;
  });
  for (auto& e : ev)
    for (hipEvent_t h : e)
      if (h) (void)hipEventDestroy(h);
  return rc;
}

int ftlm_hip_apply_host(const ftlm_op_t* op, const double* x, double* y) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && x != nullptr && y != nullptr, FTLM_EINVAL, "null argument");
    FTLM_REQUIRE(op->kind != kKStack, FTLM_EINVAL, "a momentum stack is applied to device blocks only");
    FTLM_HIP_CHECK(hipSetDevice(op->device));
    apply_host_staged(op, reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y));
  });
}

}  // extern "C"
