// batched.hip -- every (sector, sample) Lanczos recurrence of a set of small sectors in one
// persistent launch.
//
// Replaces the loops over sectors and random starts around lanczos_tridiagonal in
// ftlm_log_partition_complex / _real (src/ftlm_thermo.cpp:150-256) and in the grand-canonical
// drivers (benchmarks/bench_ftlm_nmu_rect.cpp:138-219).  A workgroup takes one (sector, sample)
// item at a time from an atomic queue (largest sectors first) and runs the whole recurrence of
// that sample:
//     w = H q;  z = <q|w>;  alpha = Re z;  w -= z q;  w -= beta_prev q_prev;
//     if |w|^2 < 1e-24: stop;  beta = |w|;  q_prev = q;  q = w / beta
// (the reference's recurrence, no reorthogonalisation).  The three vectors live in LDS for
// complex64 storage (3 * dim * 8 bytes, dim <= 6144) or in a per-workgroup slice of global memory
// for complex128; H*v is the pull form over the sector's hop tables with FP64 accumulation
// (diagonal + down bonds, then up bonds, as the one-pass sector kernel orders them), and z, |w|^2 and
// every scalar are FP64.  alpha/beta are written step-major, [k][item], for the device quadrature.

#include <algorithm>
#include <memory>
#include <vector>

#include "common.hpp"

namespace ftlm_hip {

constexpr int kBatchBlock = 512;
constexpr int kBatchMaxDimLds = 6144;  // 3 complex64 vectors in 144 KB of LDS

struct BatchSector {
  SectorView S;
  int M;            // recurrence length of this sector
  int64_t q0_off;   // start vectors of sample r at q0[q0_off + r * dim] (host-generated starts)
};

struct BatchItem {
  int sector;
  int sample;
  int out;  // output slot: sector_index * R + sample
};

namespace {

__device__ __forceinline__ uint64_t bmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kBatchBlock / 64; ++w) s += red[w];
  return s;
}

__device__ __forceinline__ double diag_of(const SectorView& S, uint32_t u, uint32_t d, int iu, int id) {
  const uint4 pl = S.planes[iu];
  const int cross = __popc(d & pl.x) + 2 * __popc(d & pl.y) + 4 * __popc(d & pl.z) + 8 * __popc(d & pl.w);
  return S.U * static_cast<double>(__popc(u & d)) + S.V * static_cast<double>(S.nuu[iu] + S.ndd[id] + cross);
}

template <typename T>
__device__ __forceinline__ void acc_hop(double& ar, double& ai, typename Cx<T>::v2 v, double cr, double ci, bool real) {
  const double vr = static_cast<double>(v.x), vi = static_cast<double>(v.y);
  ar = fma(cr, vr, ar);
  ai = fma(cr, vi, ai);
  if (!real) {
    ar = fma(-ci, vi, ar);
    ai = fma(ci, vr, ai);
  }
}

template <typename T, bool IN_LDS>
__global__ __launch_bounds__(kBatchBlock) void k_lanczos_batched(
    const BatchSector* __restrict__ secs, const BatchItem* __restrict__ items, int n_items, int* __restrict__ queue,
    int start, uint64_t seed_phase, unsigned seed, const double2* __restrict__ q0, typename Cx<T>::v2* __restrict__ gvec,
    int64_t gstride, int n_out, int force_real, double* __restrict__ alpha, double* __restrict__ beta,
    int* __restrict__ used, int* __restrict__ alive_out) {
  using V2 = typename Cx<T>::v2;
  extern __shared__ unsigned char lds_raw[];
  __shared__ double red[kBatchBlock / 64];
  __shared__ int s_item;
  __shared__ double s_scal[4];
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= n_items) break;
    const BatchItem item = items[it];
    const BatchSector& B = secs[item.sector];
    const SectorView& S = B.S;
    const int dim = static_cast<int>(S.dim);
    const int Ddn = S.Ddn;
    const bool real = force_real != 0;
    V2* base = IN_LDS ? reinterpret_cast<V2*>(lds_raw) : gvec + static_cast<int64_t>(blockIdx.x) * gstride;
    V2* q = base;
    V2* qp = base + dim;
    V2* w = base + 2 * static_cast<int64_t>(dim);
    const int o = item.out;
    // ---- start vector, normalised (src/lanczos.cpp:140-175; e_0 when |q| < 1e-18)
    double nn = 0.0;
    for (int i = threadIdx.x; i < dim; i += kBatchBlock) {
      double vr, vi;
      if (start == FTLM_START_PHASE) {
        const uint64_t sid = static_cast<uint64_t>(seed + static_cast<unsigned>(item.sample) * 100003u);
        const uint64_t h = bmix64(bmix64(seed_phase * 0x9e3779b97f4a7c15ull + sid) + static_cast<uint64_t>(i));
        const double th = 6.283185307179586 * static_cast<double>(h >> 11) * 0x1.0p-53;
        sincos(th, &vi, &vr);
      } else {
        const double2 v = q0[B.q0_off + static_cast<int64_t>(item.sample) * dim + i];
        vr = v.x;
        vi = v.y;
      }
      V2 s;
      s.x = static_cast<T>(vr);
      s.y = static_cast<T>(vi);
      q[i] = s;
      qp[i].x = qp[i].y = static_cast<T>(0);
      nn += static_cast<double>(s.x) * static_cast<double>(s.x) + static_cast<double>(s.y) * static_cast<double>(s.y);
    }
    nn = block_sum(nn, red);
    const double n0 = sqrt(nn);
    for (int i = threadIdx.x; i < dim; i += kBatchBlock) {
      V2 s = q[i];
      if (n0 < 1e-18) {
        s.x = static_cast<T>(i == 0 ? 1.0 : 0.0);
        s.y = static_cast<T>(0);
      } else {
        s.x = static_cast<T>(static_cast<double>(s.x) / n0);
        s.y = static_cast<T>(static_cast<double>(s.y) / n0);
      }
      q[i] = s;
    }
    __syncthreads();
    double bprev = 0.0;
    int nused = 0, alive = 1;
    for (int k = 0; k < B.M; ++k) {
      // ---- w = H q
      double zr = 0.0, zi = 0.0;
      for (int i = threadIdx.x; i < dim; i += kBatchBlock) {
        const int iu = i / Ddn, id = i - iu * Ddn;
        const uint32_t u = S.up_mask[iu], d = S.dn_mask[id];
        const V2 qi = q[i];
        const double dg = diag_of(S, u, d, iu, id);
        double ar = dg * static_cast<double>(qi.x), ai = dg * static_cast<double>(qi.y);
        const int rowb = iu * Ddn;
        for (int e = S.off_dn[id]; e < S.off_dn[id + 1]; ++e) {
          const uint32_t en = S.ent_dn[e];
          const uint32_t b = (en >> kTargetBits) & kBondMask;
          const bool neg = (((en >> kParityBit) & 1u) ^ (__popc(u & S.xmask_dn[b]) & 1u)) != 0;
          const double2 c = S.coef_dn[b];
          acc_hop<T>(ar, ai, q[rowb + static_cast<int>(en & kTargetMask)], neg ? -c.x : c.x, neg ? -c.y : c.y, real);
        }
        double br = 0.0, bi = 0.0;
        for (int e = S.off_up[iu]; e < S.off_up[iu + 1]; ++e) {
          const uint32_t en = S.ent_up[e];
          const uint32_t b = (en >> kTargetBits) & kBondMask;
          const bool neg = (((en >> kParityBit) & 1u) ^ (__popc(d & S.xmask_up[b]) & 1u)) != 0;
          const double2 c = S.coef_up[b];
          acc_hop<T>(br, bi, q[static_cast<int>(en & kTargetMask) * Ddn + id], neg ? -c.x : c.x, neg ? -c.y : c.y,
                     real);
        }
        V2 hv;
        hv.x = static_cast<T>(ar + br);
        hv.y = static_cast<T>(ai + bi);
        w[i] = hv;
        const double hr = static_cast<double>(hv.x), hi = static_cast<double>(hv.y);
        const double qr = static_cast<double>(qi.x), qim = static_cast<double>(qi.y);
        zr += qr * hr + qim * hi;  // <q|w> = sum conj(q) w
        zi += qr * hi - qim * hr;
      }
      zr = block_sum(zr, red);
      zi = block_sum(zi, red);
      if (threadIdx.x == 0) alpha[static_cast<int64_t>(k) * n_out + o] = zr;
      // ---- w -= z q; w -= beta_prev q_prev; |w|^2
      double nw = 0.0;
      for (int i = threadIdx.x; i < dim; i += kBatchBlock) {
        const V2 a = q[i], c = w[i];
        double wr = static_cast<double>(c.x), wi = static_cast<double>(c.y);
        const double xr = static_cast<double>(a.x), xi = static_cast<double>(a.y);
        wr -= zr * xr - zi * xi;
        wi -= zr * xi + zi * xr;
        if (k > 0) {
          const V2 b = qp[i];
          wr -= bprev * static_cast<double>(b.x);
          wi -= bprev * static_cast<double>(b.y);
        }
        V2 out;
        out.x = static_cast<T>(wr);
        out.y = static_cast<T>(wi);
        w[i] = out;
        nw += static_cast<double>(out.x) * static_cast<double>(out.x) + static_cast<double>(out.y) * static_cast<double>(out.y);
      }
      nw = block_sum(nw, red);
      nused = k + 1;
      if (nw < 1e-24) {
        alive = 0;
        break;
      }
      const double bk = sqrt(nw);
      if (threadIdx.x == 0) beta[static_cast<int64_t>(k) * n_out + o] = bk;
      bprev = bk;
      // ---- q_prev = q; q = w / beta  (rotate the three buffers)
      for (int i = threadIdx.x; i < dim; i += kBatchBlock) {
        const V2 c = w[i];
        V2 s;
        s.x = static_cast<T>(static_cast<double>(c.x) / bk);
        s.y = static_cast<T>(static_cast<double>(c.y) / bk);
        w[i] = s;
      }
      V2* t = qp;
      qp = q;
      q = w;
      w = t;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      used[o] = nused;
      alive_out[o] = alive;
    }
    (void)s_scal;
  }
}

}  // namespace

// Runs the recurrences of all (sector, sample) pairs of `secs_host` in one launch.  alpha/beta:
// device [Mmax][n_out]; used/alive: device [n_out] (n_out = n_sectors * R).
void lanczos_batched(const std::vector<const Sector*>& sectors, const std::vector<int>& Ms, int R, unsigned seed,
                     int start, int dtype, bool force_real, const std::vector<std::vector<double2>>* q0_host,
                     double* d_alpha, double* d_beta, int* d_used, int* d_alive, hipStream_t st) {
  const int ns = static_cast<int>(sectors.size());
  const int n_out = ns * R;
  std::vector<BatchSector> hs(static_cast<size_t>(ns));
  int maxdim = 0;
  int64_t q0_total = 0;
  std::vector<BatchItem> items;
  for (int s = 0; s < ns; ++s) {
    hs[s].S = sectors[s]->view();
    hs[s].M = Ms[s];
    hs[s].q0_off = q0_total;
    if (q0_host) q0_total += static_cast<int64_t>(R) * sectors[s]->dim;
    maxdim = std::max<int>(maxdim, static_cast<int>(sectors[s]->dim));
    if (Ms[s] > 0)
      for (int r = 0; r < R; ++r) items.push_back(BatchItem{s, r, s * R + r});
  }
  std::stable_sort(items.begin(), items.end(), [&](const BatchItem& a, const BatchItem& b) {
    return sectors[a.sector]->dim > sectors[b.sector]->dim;
  });
  if (items.empty()) return;
  DevBuf<BatchSector> dsecs(hs.size());
  dsecs.upload(hs.data(), hs.size(), st);
  DevBuf<BatchItem> ditems(items.size());
  ditems.upload(items.data(), items.size(), st);
  DevBuf<double2> dq0(static_cast<size_t>(std::max<int64_t>(q0_total, 1)));
  if (q0_host) {
    std::vector<double2> flat;
    flat.reserve(static_cast<size_t>(q0_total));
    for (const auto& v : *q0_host) flat.insert(flat.end(), v.begin(), v.end());
    dq0.upload(flat.data(), flat.size(), st);
  }
  DevBuf<int> queue(1);
  FTLM_HIP_CHECK(hipMemsetAsync(queue.p, 0, sizeof(int), st));
  int dev = 0, ncu = 0;
  FTLM_HIP_CHECK(hipGetDevice(&dev));
  FTLM_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const uint64_t seed_phase = 0x5eedull;
  int lds_max = 0;
  FTLM_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t lds_need = static_cast<size_t>(3) * maxdim * sizeof(float2) + 1024;
  const bool lds = dtype == FTLM_C64 && maxdim <= kBatchMaxDimLds && lds_need <= static_cast<size_t>(lds_max);
  if (lds) {
    const int grid = std::min<int>(static_cast<int>(items.size()), ncu);
    const size_t shmem = static_cast<size_t>(3) * maxdim * sizeof(float2);
    hipLaunchKernelGGL((k_lanczos_batched<float, true>), dim3(grid), dim3(kBatchBlock), shmem, st, dsecs.p, ditems.p,
                       static_cast<int>(items.size()), queue.p, start, seed_phase, seed, dq0.p, nullptr,
                       static_cast<int64_t>(0), n_out, force_real ? 1 : 0, d_alpha, d_beta, d_used, d_alive);
  } else {
    const int grid = std::min<int>(static_cast<int>(items.size()), 2 * ncu);
    const int64_t gstride = 3 * static_cast<int64_t>(maxdim);
    if (dtype == FTLM_C64) {
      DevBuf<float2> g(static_cast<size_t>(gstride) * grid);
      hipLaunchKernelGGL((k_lanczos_batched<float, false>), dim3(grid), dim3(kBatchBlock), 0, st, dsecs.p, ditems.p,
                         static_cast<int>(items.size()), queue.p, start, seed_phase, seed, dq0.p, g.p, gstride, n_out,
                         force_real ? 1 : 0, d_alpha, d_beta, d_used, d_alive);
      FTLM_HIP_CHECK(hipStreamSynchronize(st));
    } else {
      DevBuf<double2> g(static_cast<size_t>(gstride) * grid);
      hipLaunchKernelGGL((k_lanczos_batched<double, false>), dim3(grid), dim3(kBatchBlock), 0, st, dsecs.p, ditems.p,
                         static_cast<int>(items.size()), queue.p, start, seed_phase, seed, dq0.p, g.p, gstride, n_out,
                         force_real ? 1 : 0, d_alpha, d_beta, d_used, d_alive);
      FTLM_HIP_CHECK(hipStreamSynchronize(st));
    }
  }
  check_launch("k_lanczos_batched");
  FTLM_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace ftlm_hip
