// sector.hip -- lattice geometry, (N_up, N_dn) sector enumeration and ranking, pull-hop tables.
//
// Replaces, on the device:
//   FockBasis ctor / enumerate_bit_combinations      src/fock_basis.cpp:7-42
//   FockBasis::index_of (hash map)                   src/fock_basis.cpp:44-51
//   the per-element hop scan + fermions_between_modes of apply_extended_hubbard
//                                                    src/hubbard_hamiltonian.cpp:24-108
// by a combinatorial-number-system unrank/rank and per-spin tables of "pull" hops (for each
// string: the strings that hop INTO it, with the bond id and the own-spin Jordan-Wigner parity).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>

#include "common.hpp"

namespace ftlm_hip {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(FTLM_EHIP, std::string(what) + " launch failed: " + hipGetErrorString(e));
}

Binom host_binomials() {
  Binom b;
  std::memset(&b, 0, sizeof(b));
  for (int a = 0; a <= kMaxSites + 1; ++a) {
    b.c[a][0] = 1;
    for (int k = 1; k <= a; ++k) b.c[a][k] = b.c[a - 1][k - 1] + (k <= a - 1 ? b.c[a - 1][k] : 0);
  }
  return b;
}

int64_t binom(int a, int b) {
  static const Binom t = host_binomials();
  if (a < 0 || b < 0 || b > a || a > kMaxSites + 1) return 0;
  return t.c[a][b];
}

// Binomials for device-side rank/unrank.
__constant__ int64_t c_binom[kMaxSites + 2][kMaxSites + 2];

namespace {

std::once_flag g_binom_once[64];

void ensure_device_binomials(int dev) {
  std::call_once(g_binom_once[dev & 63], [] {
    const Binom b = host_binomials();
    FTLM_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), &b.c[0][0], sizeof(b.c)));
  });
}

// Lexicographic rank of a k-subset of [0, n) given as a bit mask (positions ascending):
//   rank = sum_i [ C(n-1-p_{i-1}, k-i+1) - C(n-p_i, k-i+1) ],  p_0 = -1
// (hockey-stick sum of the subsets that agree on the first i-1 positions and have a smaller
// i-th position).  Matches the enumeration order of src/fock_basis.cpp:7-15.
__device__ __forceinline__ int64_t lex_rank(uint32_t m, int n, int k) {
  int64_t r = 0;
  int prev = -1, i = 1;
  while (m) {
    const int p = __ffs(m) - 1;
    m &= m - 1;
    r += c_binom[n - 1 - prev][k - i + 1] - c_binom[n - p][k - i + 1];
    prev = p;
    ++i;
  }
  return r;
}

// Inverse of lex_rank.
__device__ __forceinline__ uint32_t lex_unrank(int64_t r, int n, int k) {
  uint32_t m = 0;
  for (int p = 0; p < n && k > 0; ++p) {
    const int64_t c = c_binom[n - 1 - p][k - 1];  // subsets whose next position is p
    if (r < c) {
      m |= 1u << p;
      --k;
    } else {
      r -= c;
    }
  }
  return m;
}

// Rank of a k-subset among all k-subsets of [0, n) in increasing order of their masks (the
// colexicographic order of the position sets): sum_i C(p_i, i), positions p_1 < p_2 < ...
__device__ __forceinline__ int64_t num_rank(uint32_t m) {
  int64_t r = 0;
  int i = 1;
  while (m) {
    const int p = __ffs(m) - 1;
    m &= m - 1;
    r += c_binom[p][i];
    ++i;
  }
  return r;
}

// Inverse of num_rank: the largest position first (the largest p with C(p, k) <= r).
__device__ __forceinline__ uint32_t num_unrank(int64_t r, int n, int k) {
  uint32_t m = 0;
  for (int p = n - 1; p >= 0 && k > 0; --p) {
    const int64_t c = c_binom[p][k];
    if (c <= r) {
      m |= 1u << p;
      r -= c;
      --k;
    }
  }
  return m;
}

// Strings of a spin in rank order: the reference's lexicographic order, or (numeric) increasing
// masks for the internal sectors of momentum stacks.
__global__ void k_unrank(uint32_t* out, int D, int n, int k, int numeric) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D) out[i] = numeric ? num_unrank(i, n, k) : lex_unrank(i, n, k);
}

struct BondDev {
  uint8_t from, to;
  uint32_t own;    // own-spin sites strictly inside the JW string
};

// Count / fill pull-hop entries: string m receives a hop along bond (from -> to) when `to` is
// occupied and `from` is empty in m; the source string is m ^ (1<<from) ^ (1<<to).
__global__ void k_hop_count(const uint32_t* masks, int D, const BondDev* bonds, int nb, int32_t* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const uint32_t m = masks[i];
  int c = 0;
  for (int b = 0; b < nb; ++b) c += ((m >> bonds[b].to) & 1u) & (((m >> bonds[b].from) & 1u) ^ 1u);
  cnt[i] = c;
}

__global__ void k_hop_fill(const uint32_t* masks, int D, const BondDev* bonds, int nb, const int32_t* off,
                           uint32_t* ent, int n, int k, int numeric) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const uint32_t m = masks[i];
  int w = off[i];
  for (int b = 0; b < nb; ++b) {
    const BondDev bd = bonds[b];
    if (!((m >> bd.to) & 1u) || ((m >> bd.from) & 1u)) continue;
    const uint32_t src = m ^ (1u << bd.from) ^ (1u << bd.to);
    const uint32_t t = static_cast<uint32_t>(numeric ? num_rank(src) : lex_rank(src, n, k));
    const uint32_t par = __popc(m & bd.own) & 1u;
    const uint32_t e = t | (static_cast<uint32_t>(b) << kTargetBits) | (par << kParityBit);
    // keep the list sorted by source rank (every H*V kernel adds a target's hops in this order)
    int q = w++;
    for (; q > off[i] && (ent[q - 1] & kTargetMask) > t; --q) ent[q] = ent[q - 1];
    ent[q] = e;
  }
}

// CSR -> slot-major copy of a hop table (entT[k * Dpad + i] = entry k of string i); the table is
// pre-filled with kPadEntry.  Slot rows are padded to a multiple of 256 strings so that any
// aligned chunk of up to 256 strings of one slot lies inside the table and loads in 16-byte
// pieces.
__global__ void k_hop_transpose(const int32_t* off, const uint32_t* ent, int D, int Dpad, uint32_t* entT,
                                int32_t* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const int o = off[i], c = off[i + 1] - o;
  cnt[i] = c;
  for (int k = 0; k < c; ++k) entT[static_cast<int64_t>(k) * Dpad + i] = ent[o + k];
}

// Per-string pieces of the diagonal  V * sum_pairs (u_i + d_i)(u_j + d_j):
//   same-spin part sum_pairs s_i s_j, and bit planes of nb(x) = sum over pair partners of s.
__global__ void k_diag_tables(const uint32_t* masks, int D, const int2* pairs, int np, int32_t* same,
                              uint4* planes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const uint32_t m = masks[i];
  int s = 0;
  uint8_t nb[kMaxSites];
  for (int q = 0; q < kMaxSites; ++q) nb[q] = 0;
  for (int p = 0; p < np; ++p) {
    const int a = pairs[p].x, b = pairs[p].y;
    const uint32_t oa = (m >> a) & 1u, ob = (m >> b) & 1u;
    s += static_cast<int>(oa & ob);
    nb[a] += static_cast<uint8_t>(ob);
    nb[b] += static_cast<uint8_t>(oa);
  }
  same[i] = s;
  if (planes) {
    uint4 pl = make_uint4(0, 0, 0, 0);
    for (int q = 0; q < kMaxSites; ++q) {
      pl.x |= static_cast<uint32_t>(nb[q] & 1) << q;
      pl.y |= static_cast<uint32_t>((nb[q] >> 1) & 1) << q;
      pl.z |= static_cast<uint32_t>((nb[q] >> 2) & 1) << q;
      pl.w |= static_cast<uint32_t>((nb[q] >> 3) & 1) << q;
    }
    planes[i] = pl;
  }
}

// Sector index of packed states (u << 16) | d of this sector (momentum Gram blocks).
__global__ void k_rank_packs(const uint32_t* packs, int64_t cnt, int n, int nu, int nd, int Ddn, int32_t* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint32_t p = packs[i];
  out[i] = static_cast<int32_t>(lex_rank(p >> 16, n, nu) * Ddn + lex_rank(p & 0xFFFFu, n, nd));
}

// index_of with the reference's key semantics: key = (up << n) | dn, decoded back into (u, d).
__global__ void k_index_of(const uint64_t* up, const uint64_t* dn, int64_t cnt, int n, int nu, int nd, int Ddn,
                           int64_t* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint64_t key = (up[i] << n) | dn[i];
  const uint64_t lowmask = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  const uint64_t u = key >> n, d = key & lowmask;
  int64_t r = -1;
  if (u <= lowmask && __popcll(u) == nu && __popcll(d) == nd)
    r = lex_rank(static_cast<uint32_t>(u), n, nu) * Ddn + lex_rank(static_cast<uint32_t>(d), n, nd);
  out[i] = r;
}

struct BondHost {
  int from, to;
  std::complex<double> coeff;
  uint32_t own, other;
};

// Merge the directed hop list per spin: equal (from, to, spin) entries are summed in list order.
void merge_bonds(const ftlm_hop_t* hops, int n_hops, int n_sites, std::vector<BondHost> out[2]) {
  std::map<std::pair<int, int>, int> idx[2];
  for (int h = 0; h < n_hops; ++h) {
    const ftlm_hop_t& e = hops[h];
    FTLM_REQUIRE(e.spin == 0 || e.spin == 1, FTLM_EINVAL, "hop spin must be 0 or 1");
    FTLM_REQUIRE(e.from >= 0 && e.from < n_sites && e.to >= 0 && e.to < n_sites, FTLM_EINVAL,
                 "hop site outside [0, n_sites)");
    if (e.from == e.to) continue;  // number operators are not hops; the reference never emits them
    auto key = std::make_pair(e.from, e.to);
    auto it = idx[e.spin].find(key);
    if (it == idx[e.spin].end()) {
      const int lo = std::min(e.from, e.to), hi = std::max(e.from, e.to);
      uint32_t inner = 0, other = 0;
      for (int s = lo + 1; s <= hi - 1; ++s) inner |= 1u << s;
      if (e.spin == 0) {
        for (int s = lo; s <= hi - 1; ++s) other |= 1u << s;  // down modes 2s+1 in (2lo, 2hi)
      } else {
        for (int s = lo + 1; s <= hi; ++s) other |= 1u << s;  // up modes 2s in (2lo+1, 2hi+1)
      }
      idx[e.spin][key] = static_cast<int>(out[e.spin].size());
      out[e.spin].push_back(BondHost{e.from, e.to, std::complex<double>(e.re, e.im), inner, other});
    } else {
      out[e.spin][static_cast<size_t>(it->second)].coeff += std::complex<double>(e.re, e.im);
    }
  }
}

int current_device() {
  int d = 0;
  FTLM_HIP_CHECK(hipGetDevice(&d));
  return d;
}

}  // namespace

int32_t* Sector::tile_counters(hipStream_t st) const {
  std::lock_guard<std::mutex> lk(ctr_mu);
  const void* key = static_cast<const void*>(st);
  size_t i = 0;
  while (i < ctr_stream.size() && ctr_stream[i] != key) ++i;
  if (i == ctr_stream.size()) {
    FTLM_REQUIRE(static_cast<int>(i) < kCtrStreams, FTLM_EUNSUPPORTED,
                 "an operator is applied from at most 64 distinct streams");
    ctr_stream.push_back(key);
  }
  return tile_ctr.p + i * kCtrSlice;
}

SectorView Sector::view() const {
  SectorView v;
  v.dim = dim;
  v.Dup = Dup;
  v.Ddn = Ddn;
  v.n_sites = n_sites;
  v.U = U;
  v.V = V;
  v.up_mask = up_mask.p;
  v.dn_mask = dn_mask.p;
  v.off_up = off[0].p;
  v.ent_up = ent[0].p;
  v.off_dn = off[1].p;
  v.ent_dn = ent[1].p;
  v.coef_up = coef[0].p;
  v.coef_dn = coef[1].p;
  v.xmask_up = xmask[0].p;
  v.xmask_dn = xmask[1].p;
  v.nuu = nuu.p;
  v.ndd = ndd.p;
  v.planes = planes.p;
  v.entT_up = entT[0].p;
  v.cnt_up = cnt[0].p;
  v.entT_dn = entT[1].p;
  v.cnt_dn = cnt[1].p;
  v.nbond_up = nbond[0];
  v.nbond_dn = nbond[1];
  v.entT_stride_up = entT_stride[0];
  v.entT_stride_dn = entT_stride[1];
  v.maxcnt_up = maxcnt[0];
  v.maxcnt_dn = maxcnt[1];
  v.st_blk_up = st_blk[0].p;
  v.st_desc_up = st_desc[0].p;
  v.st_ent_up = st_ent[0].p;
  v.st_blk_dn = st_blk[1].p;
  v.st_desc_dn = st_desc[1].p;
  v.st_ent_dn = st_ent[1].p;
  v.line_base = nullptr;
  v.line_umask = nullptr;
  return v;
}

namespace {

// Step table of one spin (layout at kStepBlock in common.hpp) from its CSR pull-hop table, whose
// per-target lists are sorted by source rank.  Each block's hops are taken in increasing source
// order; a step grows one source at a time and is closed before the source that would make it
// span more than kStepSub sources or give some wavefront more than kStepRows rows.  One source
// adds at most one row per group (a target has one hop per source), so at most kStepRounds rows
// per wavefront, and every step holds at least one source.
void build_step_table(Sector* S, int sp, int D, const std::vector<int32_t>& off, const std::vector<uint32_t>& ent) {
  const int nblk = (D + kStepBlock - 1) / kStepBlock;
  constexpr int kGroups = kStepBlock / kStepGroup;  // 128: group g = j * 16 + w
  const uint32_t pad = kStepZero * 8u << kStepPosBits;
  std::vector<int32_t> blk(static_cast<size_t>(nblk) + 1, 0);
  std::vector<uint4> desc;
  std::vector<uint32_t> rows;
  std::vector<uint64_t> trip;                          // source << 32 | local target << 16 | coefficient index
  std::vector<std::vector<uint32_t>> lst(kStepBlock);  // the step's entries of each target
  std::vector<int> touched;
  int K[kGroups], need[kGroups], wrows[kStepWaves];
  for (int b = 0; b < nblk; ++b) {
    const int t0 = b * kStepBlock, t1 = std::min(D, t0 + kStepBlock);
    trip.clear();
    for (int t = t0; t < t1; ++t)
      for (int k = off[t]; k < off[t + 1]; ++k) {
        const uint32_t e = ent[static_cast<size_t>(k)];
        const uint32_t bond = (e >> kTargetBits) & kBondMask, par = (e >> kParityBit) & 1u;
        trip.push_back(static_cast<uint64_t>(e & kTargetMask) << 32 | static_cast<uint64_t>(t - t0) << 16 |
                       (bond | par << 8));
      }
    std::sort(trip.begin(), trip.end());
    size_t i = 0;
    while (i < trip.size()) {
      const uint32_t xoff = static_cast<uint32_t>(trip[i] >> 32);
      uint32_t last = xoff;
      std::fill(K, K + kGroups, 0);
      std::fill(wrows, wrows + kStepWaves, 0);
      for (int t : touched) lst[static_cast<size_t>(t)].clear();
      touched.clear();
      while (i < trip.size()) {
        const uint32_t src = static_cast<uint32_t>(trip[i] >> 32);
        if (src - xoff >= static_cast<uint32_t>(kStepSub)) break;
        size_t j = i;
        while (j < trip.size() && static_cast<uint32_t>(trip[j] >> 32) == src) ++j;
        std::copy(K, K + kGroups, need);
        for (size_t q = i; q < j; ++q) {
          const int t = static_cast<int>((trip[q] >> 16) & 0xffffu), g = t / kStepGroup;
          need[g] = std::max(need[g], static_cast<int>(lst[static_cast<size_t>(t)].size()) + 1);
        }
        bool fits = true;
        int nr[kStepWaves];
        std::copy(wrows, wrows + kStepWaves, nr);
        for (int g = 0; g < kGroups; ++g) nr[g % kStepWaves] += need[g] - K[g];
        for (int w = 0; w < kStepWaves; ++w) fits = fits && nr[w] <= kStepRows;
        if (!fits) break;  // the step is not empty here (see above)
        for (size_t q = i; q < j; ++q) {
          const int t = static_cast<int>((trip[q] >> 16) & 0xffffu);
          if (lst[static_cast<size_t>(t)].empty()) touched.push_back(t);
          lst[static_cast<size_t>(t)].push_back((src - xoff) * 128u |
                                                static_cast<uint32_t>(trip[q] & 0xffffu) * 8u << kStepPosBits);
        }
        std::copy(need, need + kGroups, K);
        std::copy(nr, nr + kStepWaves, wrows);
        last = src;
        i = j;
      }
      const uint32_t nx = last - xoff + 1;
      for (int w = 0; w < kStepWaves; ++w) {
        const size_t row0 = rows.size() / kStepGroup;
        FTLM_REQUIRE(row0 < (size_t{1} << 26), FTLM_EUNSUPPORTED, "step hop table too large");
        uint32_t cnt[2] = {0u, 0u};
        for (int j = 0; j < kStepRounds; ++j) {
          const int g = j * kStepWaves + w;
          cnt[j / 4] |= static_cast<uint32_t>(K[g]) << (8 * (j % 4));
          for (int k = 0; k < K[g]; ++k)
            for (int q = 0; q < kStepGroup; ++q) {
              const auto& L = lst[static_cast<size_t>(g * kStepGroup + q)];
              rows.push_back(k < static_cast<int>(L.size()) ? L[static_cast<size_t>(k)] : pad);
            }
        }
        desc.push_back(make_uint4(xoff | nx << 20, static_cast<uint32_t>(row0) | static_cast<uint32_t>(wrows[w]) << 26,
                                  cnt[0], cnt[1]));
      }
    }
    blk[static_cast<size_t>(b) + 1] = static_cast<int32_t>(desc.size() / kStepWaves);
  }
  S->st_blk[sp].alloc(blk.size());
  S->st_blk[sp].upload(blk);
  S->st_desc[sp].alloc(std::max<size_t>(desc.size(), 1));
  S->st_desc[sp].upload(desc);
  S->st_ent[sp].alloc(std::max<size_t>(rows.size(), 1));
  S->st_ent[sp].upload(rows);
  FTLM_HIP_CHECK(hipDeviceSynchronize());
}

void build_sector(Sector* S, const ftlm_hop_t* hops, int n_hops, const ftlm_pair_t* pairs, int n_pairs) {
  const int n = S->n_sites;
  S->device = current_device();
  ensure_device_binomials(S->device);
  const int64_t Du = binom(n, S->n_up), Dd = binom(n, S->n_dn);
  FTLM_REQUIRE(Du < (1 << kTargetBits) && Dd < (1 << kTargetBits), FTLM_EUNSUPPORTED,
               "per-spin string count must be < 2^20 on the device path");
  S->Dup = static_cast<int>(Du);
  S->Ddn = static_cast<int>(Dd);
  S->dim = Du * Dd;

  // strings (K1: device unrank)
  S->up_mask.alloc(static_cast<size_t>(Du));
  S->dn_mask.alloc(static_cast<size_t>(Dd));
  const int numeric = S->numeric_order ? 1 : 0;
  hipLaunchKernelGGL(k_unrank, dim3(grid_for(Du, 256)), dim3(256), 0, 0, S->up_mask.p, S->Dup, n, S->n_up, numeric);
  hipLaunchKernelGGL(k_unrank, dim3(grid_for(Dd, 256)), dim3(256), 0, 0, S->dn_mask.p, S->Ddn, n, S->n_dn, numeric);
  check_launch("k_unrank");

  // bonds
  std::vector<BondHost> bonds[2];
  merge_bonds(hops, n_hops, n, bonds);
  S->real_h = true;
  for (int sp = 0; sp < 2; ++sp) {
    const int nb = static_cast<int>(bonds[sp].size());
    FTLM_REQUIRE(nb < static_cast<int>(kZeroBond), FTLM_EUNSUPPORTED,
                 "more than 510 distinct directed bonds per spin");
    S->nbond[sp] = nb;
    std::vector<BondDev> bd(static_cast<size_t>(std::max(nb, 1)));
    std::vector<double2> cf(static_cast<size_t>(std::max(nb, 1)));
    std::vector<uint32_t> xm(static_cast<size_t>(std::max(nb, 1)));
    for (int b = 0; b < nb; ++b) {
      const BondHost& h = bonds[sp][static_cast<size_t>(b)];
      bd[b] = BondDev{static_cast<uint8_t>(h.from), static_cast<uint8_t>(h.to), h.own};
      cf[b] = make_double2(h.coeff.real(), h.coeff.imag());
      xm[b] = h.other;
      S->bond_key[sp].emplace_back(h.from, h.to);
      S->bond_coef[sp].push_back(h.coeff);
      if (h.coeff.imag() != 0.0) S->real_h = false;
    }
    S->coef[sp].alloc(cf.size());
    S->coef[sp].upload(cf);
    S->xmask[sp].alloc(xm.size());
    S->xmask[sp].upload(xm);
    DevBuf<BondDev> dbonds(bd.size());
    dbonds.upload(bd);

    const int D = sp == 0 ? S->Dup : S->Ddn;
    const int k = sp == 0 ? S->n_up : S->n_dn;
    const uint32_t* masks = sp == 0 ? S->up_mask.p : S->dn_mask.p;
    DevBuf<int32_t> cnt(static_cast<size_t>(D));
    hipLaunchKernelGGL(k_hop_count, dim3(grid_for(D, 256)), dim3(256), 0, 0, masks, D, dbonds.p, nb, cnt.p);
    check_launch("k_hop_count");
    std::vector<int32_t> hc(static_cast<size_t>(D)), ho(static_cast<size_t>(D) + 1, 0);
    FTLM_HIP_CHECK(hipMemcpy(hc.data(), cnt.p, sizeof(int32_t) * D, hipMemcpyDeviceToHost));
    int64_t acc = 0;
    for (int i = 0; i < D; ++i) {
      ho[i] = static_cast<int32_t>(acc);
      acc += hc[i];
    }
    FTLM_REQUIRE(acc < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "hop table too large");
    ho[D] = static_cast<int32_t>(acc);
    S->nnz[sp] = acc;
    S->off[sp].alloc(ho.size());
    S->off[sp].upload(ho);
    S->ent[sp].alloc(static_cast<size_t>(std::max<int64_t>(acc, 1)));
    hipLaunchKernelGGL(k_hop_fill, dim3(grid_for(D, 256)), dim3(256), 0, 0, masks, D, dbonds.p, nb, S->off[sp].p,
                       S->ent[sp].p, n, k, numeric);
    check_launch("k_hop_fill");
    int mc = 0;
    for (int i = 0; i < D; ++i) mc = std::max(mc, hc[i]);
    mc = std::max(kSlotPad, (mc + kSlotPad - 1) / kSlotPad * kSlotPad);
    S->maxcnt[sp] = mc;
    const int Dpad = (D + 255) / 256 * 256;
    S->entT_stride[sp] = Dpad;
    S->entT[sp].alloc(static_cast<size_t>(mc) * Dpad);
    FTLM_HIP_CHECK(hipMemsetD32(reinterpret_cast<hipDeviceptr_t>(S->entT[sp].p), static_cast<int>(kPadEntry),
                                S->entT[sp].n));
    S->cnt[sp].alloc(static_cast<size_t>(D));
    hipLaunchKernelGGL(k_hop_transpose, dim3(grid_for(D, 256)), dim3(256), 0, 0, S->off[sp].p, S->ent[sp].p, D, Dpad,
                       S->entT[sp].p, S->cnt[sp].p);
    check_launch("k_hop_transpose");
    FTLM_HIP_CHECK(hipDeviceSynchronize());
    // step tables of the LDS-staged passes (sectors whose strings reach the two-pass size)
    if (S->Ddn >= 128 && S->Dup >= 128 && nb <= kStepMaxBonds) {
      std::vector<uint32_t> he(static_cast<size_t>(acc));
      if (acc) FTLM_HIP_CHECK(hipMemcpy(he.data(), S->ent[sp].p, sizeof(uint32_t) * acc, hipMemcpyDeviceToHost));
      build_step_table(S, sp, D, ho, he);
    }
  }

  // diagonal tables
  std::vector<int2> pr(static_cast<size_t>(std::max(n_pairs, 1)));
  for (int p = 0; p < n_pairs; ++p) {
    FTLM_REQUIRE(pairs[p].i >= 0 && pairs[p].i < n && pairs[p].j >= 0 && pairs[p].j < n, FTLM_EINVAL,
                 "pair site outside [0, n_sites)");
    pr[p] = make_int2(pairs[p].i, pairs[p].j);
    S->pair_list.emplace_back(std::min(pairs[p].i, pairs[p].j), std::max(pairs[p].i, pairs[p].j));
  }
  std::sort(S->pair_list.begin(), S->pair_list.end());
  {
    std::vector<int> deg(static_cast<size_t>(n), 0);
    for (int p = 0; p < n_pairs; ++p) {
      ++deg[pairs[p].i];
      ++deg[pairs[p].j];
    }
    for (int s = 0; s < n; ++s)
      FTLM_REQUIRE(deg[s] <= 15, FTLM_EUNSUPPORTED, "a site appears in more than 15 NN pairs");
  }
  S->tile_ctr.alloc(static_cast<size_t>(Sector::kCtrStreams) * Sector::kCtrSlice);
  FTLM_HIP_CHECK(hipMemset(S->tile_ctr.p, 0, S->tile_ctr.n * sizeof(int32_t)));
  DevBuf<int2> dpairs(pr.size());
  dpairs.upload(pr);
  S->nuu.alloc(static_cast<size_t>(S->Dup));
  S->ndd.alloc(static_cast<size_t>(S->Ddn));
  S->planes.alloc(static_cast<size_t>(S->Dup));
  hipLaunchKernelGGL(k_diag_tables, dim3(grid_for(S->Dup, 128)), dim3(128), 0, 0, S->up_mask.p, S->Dup, dpairs.p,
                     n_pairs, S->nuu.p, S->planes.p);
  hipLaunchKernelGGL(k_diag_tables, dim3(grid_for(S->Ddn, 128)), dim3(128), 0, 0, S->dn_mask.p, S->Ddn, dpairs.p,
                     n_pairs, S->ndd.p, static_cast<uint4*>(nullptr));
  check_launch("k_diag_tables");
  FTLM_HIP_CHECK(hipDeviceSynchronize());
}

void geometry(const ftlm_params_t& p, std::vector<ftlm_hop_t>* hops, std::vector<ftlm_pair_t>* pairs) {
  // build_hubbard_geometry (src/lattice.cpp:38-110): per site (row-major), east then north
  // neighbour with their NN pair, then the four diagonals (+,+) (+,-) (-,+) (-,-); each bond adds
  // both directions for both spins; Peierls phase exp(i(wx phi_x + wy phi_y)) on wrapped bonds.
  hops->clear();
  pairs->clear();
  auto bond = [&](int a, int b, double wphase, double ts) {
    if (std::abs(ts) < 1e-15) return;
    const std::complex<double> ph = std::polar(1.0, wphase);
    const std::complex<double> f = -ts * ph, r = -ts * std::conj(ph);
    for (int s = 0; s < 2; ++s) {
      hops->push_back(ftlm_hop_t{a, b, s, f.real(), f.imag()});
      hops->push_back(ftlm_hop_t{b, a, s, r.real(), r.imag()});
    }
  };
  auto pair = [&](int a, int b) {
    if (a == b) return;
    pairs->push_back(ftlm_pair_t{std::min(a, b), std::max(a, b)});
  };
  const int Lx = p.Lx, Ly = p.Ly;
  for (int y = 0; y < Ly; ++y)
    for (int x = 0; x < Lx; ++x) {
      const int i = y * Lx + x;
      if (Lx > 1) {
        const int wx = (x + 1 >= Lx) ? 1 : 0;
        const int j = y * Lx + (x + 1) % Lx;
        if (i != j) {
          pair(i, j);
          bond(i, j, static_cast<double>(wx) * p.phi_x + 0.0 * p.phi_y, p.t);
        }
      }
      if (Ly > 1) {
        const int wy = (y + 1 >= Ly) ? 1 : 0;
        const int j = ((y + 1) % Ly) * Lx + x;
        if (i != j) {
          pair(i, j);
          bond(i, j, 0.0 * p.phi_x + static_cast<double>(wy) * p.phi_y, p.t);
        }
      }
      static const int dirs[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
      for (const auto& dv : dirs) {
        int nx = x + dv[0], ny = y + dv[1], wx = 0, wy = 0;
        if (nx >= Lx) { nx -= Lx; wx = 1; } else if (nx < 0) { nx += Lx; wx = -1; }
        if (ny >= Ly) { ny -= Ly; wy = 1; } else if (ny < 0) { ny += Ly; wy = -1; }
        const int j = ny * Lx + nx;
        if (i != j) bond(i, j, static_cast<double>(wx) * p.phi_x + static_cast<double>(wy) * p.phi_y, p.tp);
      }
    }
}

}  // namespace

std::unique_ptr<Sector> sector_create_numeric(int n_sites, int n_up, int n_dn, const ftlm_hop_t* hops, int n_hops,
                                              const ftlm_pair_t* pairs, int n_pairs, double U, double V) {
  FTLM_REQUIRE(n_sites >= 1 && n_sites <= kMaxSites && n_up >= 0 && n_dn >= 0 && n_up <= n_sites &&
                   n_dn <= n_sites,
               FTLM_EINVAL, "sector out of range");
  auto S = std::make_unique<Sector>();
  S->kind = kSector;
  S->n_sites = n_sites;
  S->n_up = n_up;
  S->n_dn = n_dn;
  S->U = U;
  S->V = V;
  S->numeric_order = true;
  build_sector(S.get(), hops, n_hops, pairs, n_pairs);
  return S;
}

}  // namespace ftlm_hip

using namespace ftlm_hip;

namespace ftlm_hip {
// Sector indices of packed states (every state must belong to the sector; its dim < 2^31).
void sector_rank_packs(const Sector* S, const uint32_t* d_packs, int64_t count, int32_t* d_out, hipStream_t st) {
  FTLM_REQUIRE(S->dim < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "sector too large for 32-bit state indices");
  ensure_device_binomials(S->device);
  if (count <= 0) return;
  hipLaunchKernelGGL(k_rank_packs, dim3(grid_for(count, 256)), dim3(256), 0, st, d_packs, count, S->n_sites, S->n_up,
                     S->n_dn, S->Ddn, d_out);
  check_launch("k_rank_packs");
}
}  // namespace ftlm_hip

extern "C" {

const char* ftlm_hip_version(void) { return "0.1.0-dev+mi355x"; }

const char* ftlm_hip_last_error(void) { return g_last_error.c_str(); }

int ftlm_hip_device_count(int* count) {
  return guarded([&] {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
  });
}

int ftlm_hip_nsites(const ftlm_params_t* p) { return p ? p->Lx * p->Ly : 0; }

// d(coefficient)/d(theta) of every entry build_hubbard_geometry emits, same traversal and order
// (so duplicate bonds merge the same way).  theta: 0 = t, 1 = t', 2 = phi_x, 3 = phi_y.  A bond
// whose own amplitude is zero is still emitted for d/dt and d/dt' (the operator it multiplies).
void geometry_derivative(const ftlm_params_t& p, int which, std::vector<ftlm_hop_t>* hops) {
  hops->clear();
  auto bond = [&](int a, int b, int wx, int wy, double ts, bool nn) {
    const double wphase = static_cast<double>(wx) * p.phi_x + static_cast<double>(wy) * p.phi_y;
    const std::complex<double> ph = std::polar(1.0, wphase);
    std::complex<double> f, r;  // d/dtheta of -ts * ph and of -ts * conj(ph)
    if (which == 0 || which == 1) {
      if ((which == 0) != nn) return;
      f = -ph;
      r = -std::conj(ph);
    } else {
      const double w = which == 2 ? wx : wy;
      if (std::abs(ts) < 1e-15 || w == 0.0) return;
      const std::complex<double> I(0.0, 1.0);
      f = -ts * (I * w) * ph;
      r = -ts * (-I * w) * std::conj(ph);
    }
    for (int s = 0; s < 2; ++s) {
      hops->push_back(ftlm_hop_t{a, b, s, f.real(), f.imag()});
      hops->push_back(ftlm_hop_t{b, a, s, r.real(), r.imag()});
    }
  };
  const int Lx = p.Lx, Ly = p.Ly;
  for (int y = 0; y < Ly; ++y)
    for (int x = 0; x < Lx; ++x) {
      const int i = y * Lx + x;
      if (Lx > 1) {
        const int wx = (x + 1 >= Lx) ? 1 : 0;
        const int j = y * Lx + (x + 1) % Lx;
        if (i != j && !(std::abs(p.t) < 1e-15 && which >= 2)) bond(i, j, wx, 0, p.t, true);
      }
      if (Ly > 1) {
        const int wy = (y + 1 >= Ly) ? 1 : 0;
        const int j = ((y + 1) % Ly) * Lx + x;
        if (i != j && !(std::abs(p.t) < 1e-15 && which >= 2)) bond(i, j, 0, wy, p.t, true);
      }
      static const int dirs[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
      for (const auto& dv : dirs) {
        int nx = x + dv[0], ny = y + dv[1], wx = 0, wy = 0;
        if (nx >= Lx) { nx -= Lx; wx = 1; } else if (nx < 0) { nx += Lx; wx = -1; }
        if (ny >= Ly) { ny -= Ly; wy = 1; } else if (ny < 0) { ny += Ly; wy = -1; }
        const int j = ny * Lx + nx;
        if (i != j) bond(i, j, wx, wy, p.tp, false);
      }
    }
}

int ftlm_hip_build_geometry(const ftlm_params_t* p, ftlm_hop_t* hops, int hop_cap, int* n_hops, ftlm_pair_t* pairs,
                            int pair_cap, int* n_pairs) {
  return guarded([&] {
    FTLM_REQUIRE(p != nullptr && n_hops != nullptr && n_pairs != nullptr, FTLM_EINVAL, "null argument");
    FTLM_REQUIRE(p->Lx >= 1 && p->Ly >= 1, FTLM_EINVAL, "Lx, Ly must be >= 1");
    std::vector<ftlm_hop_t> h;
    std::vector<ftlm_pair_t> q;
    geometry(*p, &h, &q);
    *n_hops = static_cast<int>(h.size());
    *n_pairs = static_cast<int>(q.size());
    FTLM_REQUIRE(static_cast<int>(h.size()) <= hop_cap && static_cast<int>(q.size()) <= pair_cap, FTLM_ERANGE,
                 "geometry buffers too small");
    if (!h.empty()) std::memcpy(hops, h.data(), h.size() * sizeof(ftlm_hop_t));
    if (!q.empty()) std::memcpy(pairs, q.data(), q.size() * sizeof(ftlm_pair_t));
  });
}

int ftlm_hip_build_geometry_derivative(const ftlm_params_t* p, int which, ftlm_hop_t* hops, int hop_cap,
                                       int* n_hops) {
  return guarded([&] {
    FTLM_REQUIRE(p != nullptr && n_hops != nullptr, FTLM_EINVAL, "null argument");
    FTLM_REQUIRE(which >= 0 && which <= 3, FTLM_EINVAL, "which must be 0 (t), 1 (t'), 2 (phi_x) or 3 (phi_y)");
    FTLM_REQUIRE(p->Lx >= 1 && p->Ly >= 1 && p->Lx * p->Ly <= kMaxSites, FTLM_EINVAL, "bad lattice");
    std::vector<ftlm_hop_t> h;
    geometry_derivative(*p, which, &h);
    *n_hops = static_cast<int>(h.size());
    FTLM_REQUIRE(static_cast<int>(h.size()) <= hop_cap, FTLM_ERANGE, "hop buffer too small");
    if (!h.empty()) std::memcpy(hops, h.data(), h.size() * sizeof(ftlm_hop_t));
  });
}

int ftlm_hip_sector_create(int n_sites, int n_up, int n_dn, const ftlm_hop_t* hops, int n_hops,
                           const ftlm_pair_t* pairs, int n_pairs, double U, double V, ftlm_op_t** out) {
  return guarded([&] {
    FTLM_REQUIRE(out != nullptr, FTLM_EINVAL, "null output handle");
    *out = nullptr;
    FTLM_REQUIRE(n_sites >= 1 && n_sites <= kMaxSites, FTLM_EINVAL,
                 "FockBasis: require 1 <= n_sites <= 32 (bit packing)");
    FTLM_REQUIRE(n_up >= 0 && n_dn >= 0 && n_up <= n_sites && n_dn <= n_sites, FTLM_EINVAL,
                 "FockBasis: particle numbers out of range");
    FTLM_REQUIRE(n_hops >= 0 && n_pairs >= 0 && (n_hops == 0 || hops) && (n_pairs == 0 || pairs), FTLM_EINVAL,
                 "null hop / pair list");
    auto S = std::make_unique<Sector>();
    S->kind = kSector;
    S->n_sites = n_sites;
    S->n_up = n_up;
    S->n_dn = n_dn;
    S->U = U;
    S->V = V;
    build_sector(S.get(), hops, n_hops, pairs, n_pairs);
    *out = S.release();
  });
}

int ftlm_hip_op_destroy(ftlm_op_t* op) {
  return guarded([&] { delete op; });
}

int ftlm_hip_op_dim(const ftlm_op_t* op, int64_t* dim, int64_t* dim_up, int64_t* dim_dn) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr, FTLM_EINVAL, "null operator");
    if (dim) *dim = op->dim;
    const Sector* S = op->kind == kSector ? static_cast<const Sector*>(op) : nullptr;
    if (dim_up) *dim_up = S ? S->Dup : 0;
    if (dim_dn) *dim_dn = S ? S->Ddn : 0;
  });
}

int ftlm_hip_sector_masks(const ftlm_op_t* op, int64_t first, int64_t count, uint64_t* up, uint64_t* dn) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && op->kind == kSector, FTLM_EINVAL, "not a full-sector operator");
    const Sector* S = static_cast<const Sector*>(op);
    FTLM_REQUIRE(first >= 0 && count >= 0 && first + count <= S->dim, FTLM_ERANGE, "FockBasis index out of range");
    if (count == 0) return;
    std::vector<uint32_t> hu(static_cast<size_t>(S->Dup)), hd(static_cast<size_t>(S->Ddn));
    FTLM_HIP_CHECK(hipMemcpy(hu.data(), S->up_mask.p, sizeof(uint32_t) * S->Dup, hipMemcpyDeviceToHost));
    FTLM_HIP_CHECK(hipMemcpy(hd.data(), S->dn_mask.p, sizeof(uint32_t) * S->Ddn, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < count; ++i) {
      const int64_t g = first + i;
      if (up) up[i] = hu[static_cast<size_t>(g / S->Ddn)];
      if (dn) dn[i] = hd[static_cast<size_t>(g % S->Ddn)];
    }
  });
}

int ftlm_hip_sector_strings(const ftlm_op_t* op, int spin, uint32_t* out) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && op->kind == kSector, FTLM_EINVAL, "not a full-sector operator");
    FTLM_REQUIRE(spin == 0 || spin == 1, FTLM_EINVAL, "spin must be 0 or 1");
    FTLM_REQUIRE(out != nullptr, FTLM_EINVAL, "null output");
    const Sector* S = static_cast<const Sector*>(op);
    const int D = spin == 0 ? S->Dup : S->Ddn;
    FTLM_HIP_CHECK(hipMemcpy(out, spin == 0 ? S->up_mask.p : S->dn_mask.p, sizeof(uint32_t) * D,
                             hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_index_of(const ftlm_op_t* op, const uint64_t* up, const uint64_t* dn, int64_t count, int64_t* out) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && op->kind == kSector, FTLM_EINVAL, "not a full-sector operator");
    const Sector* S = static_cast<const Sector*>(op);
    if (count <= 0) return;
    FTLM_HIP_CHECK(hipSetDevice(S->device));
    ensure_device_binomials(S->device);
    DevBuf<uint64_t> du(static_cast<size_t>(count)), dd(static_cast<size_t>(count));
    DevBuf<int64_t> dr(static_cast<size_t>(count));
    du.upload(up, static_cast<size_t>(count));
    dd.upload(dn, static_cast<size_t>(count));
    hipLaunchKernelGGL(k_index_of, dim3(grid_for(count, 256)), dim3(256), 0, 0, du.p, dd.p, count, S->n_sites,
                       S->n_up, S->n_dn, S->Ddn, dr.p);
    check_launch("k_index_of");
    FTLM_HIP_CHECK(hipMemcpy(out, dr.p, sizeof(int64_t) * count, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
