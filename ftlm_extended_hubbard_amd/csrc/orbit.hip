// orbit.hip -- translation orbits and momentum (k) blocks on the device (gfx950).
//
// Replaces, on the device:
//   translate_occupation_bits_rect        include/ftlm/symmetry/occup_bits_translate.hpp:10-25
//   fermionic_translation_sign_rect       include/ftlm/symmetry/fermionic_translation.hpp:32-54
//   canonical_raw_state_rect / translation_stabilizer_rect   include/ftlm/symmetry/orbit.hpp:93-155
//   translation_bloch_phase               include/ftlm/symmetry/momentum_sector.hpp:57-63
//   OrbitKBlockBasis::build / reduced_index_of_rep / map_raw_to_rep
//                                         src/orbit_k_block_basis.cpp:56-149
//   apply_orbit_k_block                   src/orbit_k_block_basis.cpp:151-220
//
// Orbit graph (shared by every momentum of one sector and Hamiltonian).  States are scanned in
// numeric (u, d) order, so the orbit representatives (minimal pack (u << 16) | d over the
// orbit) come out sorted by pack, which is the reference's row order of every k-block.  For
// each representative we keep its stabiliser as a 16-bit mask over translations
// t = dy * Lx + dx and the fermionic sign of each stabiliser element.  The Hamiltonian is
// stored once, K-independent, in pull form: the reference's push form (column beta, hop
// rep(beta) -> s', s' = T_R rep(alpha)) is generated per column and stably sorted by target
// orbit alpha, so row alpha lists its (beta, hop) terms in the reference's accumulation order.
// An entry holds beta, the translation R, the sign jw(hop) * eta(alpha, R) and the merged
// bond; the diagonal is an entry of its own at column alpha.
//
// A momentum K then only needs A_K(alpha) = sum_{S in Stab} eta_S chi*_K(S) per orbit (the
// rows with |A_K| > tol, the map orbit -> row, sqrt|A_K|) and its 16 Bloch phases:
//   y[a] = sum_entries c_bond * sign * sqrt(|A_a| / |A_b|) * chi*_K(R) * x[b]
// with the unconjugated phase of translation_bloch_phase, as the reference applies it.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "common.hpp"

namespace ftlm_hip {

constexpr int kMaxOrbitSites = 16;   // RawState packing (u << 16) | d
constexpr uint32_t kColMask = 0xFFFFFFu;
constexpr uint32_t kDiagFlag = 1u << 29;
constexpr int kStabShift = 9;  // entry .y: bond (bits 0..8), |Stab_row| - 1 (9..12), |Stab_col| - 1 (13..16)

// Lattice translation tables (kernel argument): tsite[t][s] = image of site s under
// translation t = dy * lx + dx.
struct LatDev {
  int lx, ly, n, nt;
  uint8_t tsite[kMaxOrbitSites][kMaxOrbitSites];
};

LatDev make_lat(int lx, int ly) {
  LatDev L{};
  L.lx = lx;
  L.ly = ly;
  L.n = lx * ly;
  L.nt = lx * ly;
  for (int dy = 0; dy < ly; ++dy)
    for (int dx = 0; dx < lx; ++dx)
      for (int s = 0; s < L.n; ++s) {
        const int x = s % lx, y = s / lx;
        L.tsite[dy * lx + dx][s] = static_cast<uint8_t>(((y + dy) % ly) * lx + (x + dx) % lx);
      }
  return L;
}

__host__ __device__ __forceinline__ uint32_t tr_bits(uint32_t m, const LatDev& L, int t) {
  uint32_t out = 0;
  while (m) {
    const int s = __builtin_ctz(m);
    m &= m - 1;
    out |= 1u << L.tsite[t][s];
  }
  return out;
}

__host__ __device__ __forceinline__ uint32_t pack_state(uint32_t u, uint32_t d) { return (u << 16) | d; }

// Parity of the inversions of the translated mode list, modes listed by site ascending with up
// (2s) before down (2s+1): the interleaved mode order itself.
__host__ __device__ __forceinline__ int tr_parity(uint32_t u, uint32_t d, const LatDev& L, int t) {
  uint64_t seen = 0;
  int inv = 0;
  for (int s = 0; s < L.n; ++s) {
    const int ts = L.tsite[t][s];
    if ((u >> s) & 1u) {
      const int m = 2 * ts;
      inv += __builtin_popcountll(seen >> (m + 1));
      seen |= uint64_t{1} << m;
    }
    if ((d >> s) & 1u) {
      const int m = 2 * ts + 1;
      inv += __builtin_popcountll(seen >> (m + 1));
      seen |= uint64_t{1} << m;
    }
  }
  return inv & 1;
}

// Canonical pack (minimum over the orbit), stabiliser mask and odd-sign mask of the stabiliser.
__host__ __device__ __forceinline__ uint32_t canon_state(uint32_t u, uint32_t d, const LatDev& L, uint32_t* stab,
                                                         uint32_t* eta) {
  const uint32_t self = pack_state(u, d);
  uint32_t best = 0xFFFFFFFFu, st = 0, et = 0;
  for (int t = 0; t < L.nt; ++t) {
    const uint32_t p = pack_state(tr_bits(u, L, t), tr_bits(d, L, t));
    best = p < best ? p : best;
    if (p == self) {
      st |= 1u << t;
      if (stab && tr_parity(u, d, L, t)) et |= 1u << t;
    }
  }
  if (stab) *stab = st;
  if (eta) *eta = et;
  return best;
}

// First translation R (dx outer, dy inner) with T_R rep = s; -1 if none.
__host__ __device__ __forceinline__ int find_translation(uint32_t rep, uint32_t s, const LatDev& L) {
  const uint32_t ru = rep >> 16, rd = rep & 0xFFFFu, su = s >> 16, sd = s & 0xFFFFu;
  for (int dx = 0; dx < L.lx; ++dx)
    for (int dy = 0; dy < L.ly; ++dy) {
      const int t = dy * L.lx + dx;
      if (tr_bits(ru, L, t) == su && tr_bits(rd, L, t) == sd) return t;
    }
  return -1;
}

__device__ __forceinline__ int64_t lower_bound_u32(const uint32_t* a, int64_t n, uint32_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Occupied modes strictly between modes ma and mb, interleaved order m = 2 * site + spin
// (fermions_between_modes, src/hubbard_hamiltonian.cpp:24-43).
__device__ __forceinline__ uint32_t interleave16(uint32_t x) {
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

__device__ __forceinline__ int modes_between(int ma, int mb, uint32_t u, uint32_t d) {
  const int lo = ma < mb ? ma : mb, hi = ma < mb ? mb : ma;
  const uint32_t occ = interleave16(u) | (interleave16(d) << 1);
  const uint32_t below_hi = hi >= 32 ? 0xFFFFFFFFu : ((1u << hi) - 1u);
  const uint32_t upto_lo = (2u << lo) - 1u;
  return __popc(occ & below_hi & ~upto_lo);
}

struct OrbBond {
  int from, to, spin;
};

// ------------------------------------------------------------------ graph kernels
__global__ void k_orb_flag(const uint32_t* upN, int Du, const uint32_t* dnN, int Dd, LatDev L, uint8_t* flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(Du) * Dd) return;
  const uint32_t u = upN[i / Dd], d = dnN[i % Dd];
  flag[i] = canon_state(u, d, L, nullptr, nullptr) == pack_state(u, d) ? 1 : 0;
}

__global__ void k_orb_reps(const int32_t* idx, int64_t nrep, const uint32_t* upN, const uint32_t* dnN, int Dd,
                           LatDev L, const int2* pairs, int npairs, double U, double V, uint32_t* pack,
                           uint16_t* stab, uint16_t* eta, double* diag) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nrep) return;
  const int64_t i = idx[g];
  const uint32_t u = upN[i / Dd], d = dnN[i % Dd];
  uint32_t st = 0, et = 0;
  canon_state(u, d, L, &st, &et);
  pack[g] = pack_state(u, d);
  stab[g] = static_cast<uint16_t>(st);
  eta[g] = static_cast<uint16_t>(et);
  // diagonal in the reference's order: U * n_double, then + V * o_i o_j pair by pair
  double dg = static_cast<double>(__popc(u & d)) * U;
  for (int p = 0; p < npairs; ++p) {
    const int oi = static_cast<int>(((u >> pairs[p].x) & 1u) + ((d >> pairs[p].x) & 1u));
    const int oj = static_cast<int>(((u >> pairs[p].y) & 1u) + ((d >> pairs[p].y) & 1u));
    dg += V * static_cast<double>(oi * oj);
  }
  diag[g] = dg;
}

__device__ __forceinline__ bool hop_applies(const OrbBond& b, uint32_t u, uint32_t d) {
  const uint32_t s = b.spin == 0 ? u : d;
  return ((s >> b.from) & 1u) && !((s >> b.to) & 1u);
}

__global__ void k_orb_count(const uint32_t* pack, int64_t nrep, const OrbBond* bonds, int nb, int32_t* cnt) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nrep) return;
  const uint32_t u = pack[g] >> 16, d = pack[g] & 0xFFFFu;
  int c = 1;  // diagonal
  for (int b = 0; b < nb; ++b) c += hop_applies(bonds[b], u, d) ? 1 : 0;
  cnt[g] = c;
}

// Push-form entries of column g: key = target orbit, value = {g | R << 24 | sign << 28, bond}.
__global__ void k_orb_fill(const uint32_t* pack, const uint16_t* stab, int64_t nrep, const OrbBond* bonds, int nb,
                           const int32_t* off, LatDev L, uint32_t* keys, uint2* vals, int* bad) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nrep) return;
  const uint32_t u = pack[g] >> 16, d = pack[g] & 0xFFFFu;
  int64_t w = off[g];
  keys[w] = static_cast<uint32_t>(g);
  vals[w] = make_uint2(static_cast<uint32_t>(g) | kDiagFlag, 0u);
  ++w;
  for (int b = 0; b < nb; ++b) {
    const OrbBond bd = bonds[b];
    if (!hop_applies(bd, u, d)) continue;
    const uint32_t flip = (1u << bd.from) ^ (1u << bd.to);
    const uint32_t u2 = bd.spin == 0 ? u ^ flip : u, d2 = bd.spin == 1 ? d ^ flip : d;
    const int jw = modes_between(2 * bd.from + bd.spin, 2 * bd.to + bd.spin, u, d) & 1;
    const uint32_t sp = pack_state(u2, d2);
    const uint32_t cp = canon_state(u2, d2, L, nullptr, nullptr);
    const int64_t a = lower_bound_u32(pack, nrep, cp);
    int t = -1;
    if (a < nrep && pack[a] == cp) t = find_translation(cp, sp, L);
    if (t < 0) {  // cannot happen for a closed sector; reported to the host
      atomicExch(bad, 1);
      keys[w] = static_cast<uint32_t>(g);
      vals[w] = make_uint2(static_cast<uint32_t>(g) | kDiagFlag, 0xFFFFFFFFu);
      ++w;
      continue;
    }
    const int eta = tr_parity(cp >> 16, cp & 0xFFFFu, L, t);
    // |Stab| of the target (row) and source (column) orbits, for sqrt(|A_a| / |A_b|): a surviving
    // orbit has |A_K| = |Stab| at every momentum (orthogonality of the stabiliser characters)
    const uint32_t sa = __popc(static_cast<uint32_t>(stab[a])) - 1u, sb = __popc(static_cast<uint32_t>(stab[g])) - 1u;
    keys[w] = static_cast<uint32_t>(a);
    vals[w] = make_uint2(static_cast<uint32_t>(g) | (static_cast<uint32_t>(t) << 24) |
                             (static_cast<uint32_t>(jw ^ eta) << 28),
                         static_cast<uint32_t>(b) | (sa << kStabShift) | (sb << (kStabShift + 4)));
    ++w;
  }
}

__global__ void k_csr_rows(const uint32_t* keys, int64_t nnz, int32_t* row_off) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  if (i == 0 || keys[i - 1] != keys[i]) row_off[keys[i]] = static_cast<int32_t>(i);
}

// ------------------------------------------------------------------ per-momentum kernels
struct Phases {
  double2 chi[kMaxOrbitSites];  // translation_bloch_phase(K, dx, dy) at t = dy * lx + dx
};

// A_K = sum over the stabiliser (t ascending) of eta_t * chi_t, as the reference accumulates it
// ((eta, 0) * chi is exact, so the sum is a plain ordered sum of +-chi).
__global__ void k_kblock_A(const uint16_t* stab, const uint16_t* eta, int64_t nrep, int nt, Phases ph, double tol,
                           double2* A, uint8_t* member) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nrep) return;
  double re = 0.0, im = 0.0;
  const uint32_t st = stab[g], et = eta[g];
  for (int t = 0; t < nt; ++t) {
    if (!((st >> t) & 1u)) continue;
    const double2 c = ph.chi[t];
    if ((et >> t) & 1u) {
      re = __dadd_rn(re, -c.x);
      im = __dadd_rn(im, -c.y);
    } else {
      re = __dadd_rn(re, c.x);
      im = __dadd_rn(im, c.y);
    }
  }
  A[g] = make_double2(re, im);
  member[g] = hypot(re, im) > tol ? 1 : 0;
}

__global__ void k_kblock_maps(const int32_t* rows, int64_t dk, const double2* A, int32_t* map, double* sq,
                              double* isq) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= dk) return;
  const int32_t g = rows[r];
  const double a = hypot(A[g].x, A[g].y);
  map[g] = static_cast<int32_t>(r);
  sq[g] = sqrt(a);
  isq[g] = 1.0 / sqrt(a);
}

// Column of every stored term in this block's row numbering (-1: the source orbit is not a row).
__global__ void k_kblock_cols(const uint2* ent, int64_t nnz, const int32_t* map, int32_t* col) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < nnz) col[e] = map[ent[e].x & kColMask];
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double* p, int64_t n, double v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Batched map_raw_to_rep / reduced_index_of_rep / translation primitives.
__global__ void k_kblock_map_states(const uint32_t* states, int64_t count, const uint32_t* pack, int64_t nrep,
                                    const int32_t* map, LatDev L, int64_t* index, int* dx, int* dy, double* eta) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t s = states[i];
  const uint32_t cp = canon_state(s >> 16, s & 0xFFFFu, L, nullptr, nullptr);
  const int64_t g = lower_bound_u32(pack, nrep, cp);
  int64_t r = -1;
  int t = -1;
  if (g < nrep && pack[g] == cp && map[g] >= 0) {
    t = find_translation(cp, s, L);
    if (t >= 0) r = map[g];
  }
  index[i] = r;
  dx[i] = r >= 0 ? t % L.lx : 0;
  dy[i] = r >= 0 ? t / L.lx : 0;
  eta[i] = r >= 0 ? (tr_parity(cp >> 16, cp & 0xFFFFu, L, t) ? -1.0 : 1.0) : 0.0;
}

__global__ void k_kblock_index_of_rep(const uint32_t* reps, int64_t count, const uint32_t* pack, int64_t nrep,
                                      const int32_t* map, int64_t* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t g = lower_bound_u32(pack, nrep, reps[i]);
  out[i] = (g < nrep && pack[g] == reps[i]) ? map[g] : -1;
}

__global__ void k_kblock_entries(const int32_t* rows, int64_t first, int64_t count, const uint32_t* pack,
                                 const uint16_t* stab, const double2* A, int nt, uint32_t* rep, int* osz, int* ssz,
                                 double2* aout) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t g = rows[first + i];
  const int s = __popc(static_cast<uint32_t>(stab[g]));
  rep[i] = pack[g];
  ssz[i] = s;
  osz[i] = nt / s;
  aout[i] = A[g];
}

__global__ void k_translate_states(const uint32_t* states, int64_t count, LatDev L, int t, uint32_t* out,
                                   double* sign) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t u = states[i] >> 16, d = states[i] & 0xFFFFu;
  if (out) out[i] = pack_state(tr_bits(u, L, t), tr_bits(d, L, t));
  if (sign) sign[i] = tr_parity(u, d, L, t) ? -1.0 : 1.0;
}

__global__ void k_canonical_states(const uint32_t* states, int64_t count, LatDev L, uint32_t* canon,
                                   int* stab_size) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t st = 0, et = 0;
  const uint32_t c = canon_state(states[i] >> 16, states[i] & 0xFFFFu, L, &st, &et);
  if (canon) canon[i] = c;
  if (stab_size) stab_size[i] = __popc(st);
}

// ------------------------------------------------------------------ apply
// One thread per (block row, group of RB vectors).  The row's terms are read in groups of
// kKbGroup with their block columns (resolved per momentum) and gathered together; a term whose
// source orbit is not a row of this block, and the padding of the last group, get coefficient 0
// and read the row's own element.  Coefficient of a hop term, in the reference's order:
//   ((c_bond * (sign * sqrt(|Stab_row| / |Stab_col|))) * chi*_K(R)).
// Rows of a grid-strided momentum-block or momentum-stack apply, dealt per XCD.  Workgroups are
// placed on the 8 XCDs round-robin (blockIdx % 8), so with a grid that is a multiple of 8, XCD x
// walks the contiguous x-th eighth of the rows: the rows in flight on one XCD are then
// (grid / 8) * rpb consecutive orbits, and the sources their down hops share (orbits with the
// same up-string representative are consecutive) stay in that XCD's L2.  Other grids walk all
// rows grid-strided.
struct RowWalk {
  int64_t first, end, stride;
};

__device__ __forceinline__ RowWalk row_walk(int64_t n, int rpb, int rl) {
  if ((gridDim.x & 7u) == 0) {
    const int64_t x = blockIdx.x & 7u, per = gridDim.x >> 3;
    const int64_t lo = n * x / 8, hi = n * (x + 1) / 8;
    return {lo + static_cast<int64_t>(blockIdx.x >> 3) * rpb + rl, hi, per * rpb};
  }
  return {static_cast<int64_t>(blockIdx.x) * rpb + rl, n, static_cast<int64_t>(gridDim.x) * rpb};
}

// Coefficient of a stored term: the diagonal entry carries the orbit's diagonal dg, a hop term
// ((c_bond * (sign * sqrt(|Stab_row| / |Stab_col|))) * chi*_K(R)) with chi the momentum's 16
// Bloch phases.  Branch-free (every term reads its three table entries, all indices in range), so
// that the LDS reads of a group of terms issue together.
__device__ __forceinline__ void kb_term_coef(const uint2 en, const double dg, const double* s_nrm,
                                             const double2* s_coef, const double2* s_chi, double* cr, double* ci) {
  const double nrm = s_nrm[(en.y >> kStabShift) & 0xFFu];
  const double w = ((en.x >> 28) & 1u) ? -nrm : nrm;
  const double2 cb = s_coef[en.y & kBondMask];
  const double2 ch = s_chi[(en.x >> 24) & 15u];
  const double hr = cb.x * w, hi = cb.y * w;
  const double hr2 = hr * ch.x - hi * ch.y;
  const double hi2 = hr * ch.y + hi * ch.x;
  const bool diag = (en.x & kDiagFlag) != 0;
  *cr = diag ? dg : hr2;
  *ci = diag ? 0.0 : hi2;
}

template <int RB>
struct KbGroup {
  static constexpr int value = RB <= 2 ? 8 : (RB <= 4 ? 4 : 2);
};

template <typename T, int RB>
__global__ __launch_bounds__(256) void k_kblock_apply(
    const int32_t* __restrict__ rows, int64_t dk, const int32_t* __restrict__ row_off, const uint2* __restrict__ ent,
    const int32_t* __restrict__ colk, const double* __restrict__ diag, const double2* __restrict__ coef, int nbond,
    Phases ph, const uint8_t* __restrict__ acc, const typename Cx<T>::v2* __restrict__ x,
    typename Cx<T>::v2* __restrict__ y, int RT, int r0, const double* __restrict__ xs, double2* __restrict__ dotp) {
  using V2 = typename Cx<T>::v2;
  constexpr int G = KbGroup<RB>::value;
  __shared__ double2 s_coef[kMaxBondsLds];
  __shared__ double2 s_chi[kMaxOrbitSites];
  __shared__ double s_nrm[kMaxOrbitSites * kMaxOrbitSites];
  for (int b = threadIdx.x; b < kMaxBondsLds; b += blockDim.x) s_coef[b] = b < nbond ? coef[b] : make_double2(0.0, 0.0);
  if (threadIdx.x < kMaxOrbitSites) s_chi[threadIdx.x] = ph.chi[threadIdx.x];
  for (int i = threadIdx.x; i < kMaxOrbitSites * kMaxOrbitSites; i += blockDim.x)
    s_nrm[i] = sqrt(static_cast<double>((i & 15) + 1) / static_cast<double>((i >> 4) + 1));
  __syncthreads();
  double dre[RB], dimg[RB];
#pragma unroll
  for (int v = 0; v < RB; ++v) dre[v] = dimg[v] = 0.0;
  const RowWalk rw = row_walk(dk, blockDim.x, threadIdx.x);
  for (int64_t r = rw.first; r < rw.end; r += rw.stride) {
    const int32_t g = rows[r];
    const int32_t e0 = row_off[g], e1 = row_off[g + 1];
    const double dg = diag[g];
    T ar[RB], ai[RB];
#pragma unroll
    for (int v = 0; v < RB; ++v) ar[v] = ai[v] = static_cast<T>(0);
    for (int32_t e = e0; e < e1; e += G) {
      uint2 en[G];
      int32_t cc[G];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const int32_t ej = min(e + j, e1 - 1);
        en[j] = ent[ej];
        const int32_t c = colk[ej];
        cc[j] = (e + j < e1 && c >= 0) ? c : -1;
      }
      V2 src[G][RB];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const V2* sp = x + static_cast<int64_t>(cc[j] >= 0 ? cc[j] : r) * RT + r0;
#pragma unroll
        for (int v = 0; v < RB; ++v) src[j][v] = sp[v];
      }
#pragma unroll
      for (int j = 0; j < G; ++j) {
        double cr, ci;
        kb_term_coef(en[j], dg, s_nrm, s_coef, s_chi, &cr, &ci);
        if (cc[j] < 0) cr = ci = 0.0;
        const T tr = static_cast<T>(cr), ti = static_cast<T>(ci);
#pragma unroll
        for (int v = 0; v < RB; ++v) {
          ar[v] = fma(tr, src[j][v].x, ar[v]);
          ai[v] = fma(tr, src[j][v].y, ai[v]);
          ar[v] = fma(-ti, src[j][v].y, ar[v]);
          ai[v] = fma(ti, src[j][v].x, ai[v]);
        }
      }
    }
    V2* dst = y + r * RT + r0;
    const V2* own = x + r * RT + r0;
    const bool add = acc != nullptr && acc[g] != 0;  // the row's down hops are in y already
#pragma unroll
    for (int v = 0; v < RB; ++v) {
      const double sx = xs ? xs[r0 + v] : 1.0;
      V2 o;
      o.x = static_cast<T>(static_cast<double>(ar[v]) * sx);
      o.y = static_cast<T>(static_cast<double>(ai[v]) * sx);
      if (add) {
        const V2 p = dst[v];
        o.x = p.x + o.x;
        o.y = p.y + o.y;
      }
      dst[v] = o;
      if (dotp) {
        const V2 q = own[v];
        const double qr = sx * static_cast<double>(q.x), qi = sx * static_cast<double>(q.y);
        dre[v] += qr * static_cast<double>(o.x) + qi * static_cast<double>(o.y);
        dimg[v] += qr * static_cast<double>(o.y) - qi * static_cast<double>(o.x);
      }
    }
  }
  if (dotp) {
    __shared__ double red[256 / 64][RB][2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < RB; ++v) {
      double a = dre[v], b = dimg[v];
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
      }
      if (lane == 0) {
        red[wv][v][0] = a;
        red[wv][v][1] = b;
      }
    }
    __syncthreads();
    if (threadIdx.x < RB) {
      double a = 0.0, b = 0.0;
      for (int w = 0; w < 256 / 64; ++w) {
        a += red[w][threadIdx.x][0];
        b += red[w][threadIdx.x][1];
      }
      dotp[static_cast<int64_t>(blockIdx.x) * RT + r0 + threadIdx.x] = make_double2(a, b);
    }
  }
}

// Wide blocks: the RB vectors of a row are spread over L lanes of 16 bytes (VPL vectors each), as
// in the full-sector kernels, so a term's gather is one contiguous RB * s run read by L lanes; a
// workgroup covers 256 / L rows and every lane keeps 8 terms in flight.
// COOP as in k_kstack_apply: lanes 0 .. G-1 of a row load the next G terms' entries and block
// columns, and the row's lanes share them by cross-lane reads (L = 8 or 16 lanes per row).
template <typename T, int RB, bool COOP>
__global__ __launch_bounds__(256) void k_kblock_apply_split(
    const int32_t* __restrict__ rows, int64_t dk, const int32_t* __restrict__ row_off, const uint2* __restrict__ ent,
    const int32_t* __restrict__ colk, const double* __restrict__ diag, const double2* __restrict__ coef, int nbond,
    Phases ph, const uint8_t* __restrict__ acc, const typename Cx<T>::v2* __restrict__ x,
    typename Cx<T>::v2* __restrict__ y, int RT, int r0, const double* __restrict__ xs, double2* __restrict__ dotp) {
  using V2 = typename Cx<T>::v2;
  constexpr int BYTES = RB * static_cast<int>(sizeof(V2));
  constexpr int L = BYTES >= 32 ? BYTES / 16 : 1;
  constexpr int VPL = RB / L;
  constexpr int RPB = 256 / L;  // rows per workgroup pass
  constexpr int G = 8;
  __shared__ double2 s_coef[kMaxBondsLds];
  __shared__ double2 s_chi[kMaxOrbitSites];
  __shared__ double s_nrm[kMaxOrbitSites * kMaxOrbitSites];
  for (int b = threadIdx.x; b < kMaxBondsLds; b += blockDim.x) s_coef[b] = b < nbond ? coef[b] : make_double2(0.0, 0.0);
  if (threadIdx.x < kMaxOrbitSites) s_chi[threadIdx.x] = ph.chi[threadIdx.x];
  for (int i = threadIdx.x; i < kMaxOrbitSites * kMaxOrbitSites; i += blockDim.x)
    s_nrm[i] = sqrt(static_cast<double>((i & 15) + 1) / static_cast<double>((i >> 4) + 1));
  __syncthreads();
  const int sub = static_cast<int>(threadIdx.x) % L, rl = static_cast<int>(threadIdx.x) / L;
  const int rr = r0 + sub * VPL;
  double dre[VPL], dimg[VPL], sxv[VPL];
#pragma unroll
  for (int v = 0; v < VPL; ++v) {
    dre[v] = dimg[v] = 0.0;
    sxv[v] = xs ? xs[rr + v] : 1.0;
  }
  const RowWalk rw = row_walk(dk, RPB, rl);
  for (int64_t r = rw.first; r < rw.end; r += rw.stride) {
    const int32_t g = rows[r];
    const int32_t e0 = row_off[g], e1 = row_off[g + 1];
    const double dg = diag[g];
    T ar[VPL], ai[VPL];
#pragma unroll
    for (int v = 0; v < VPL; ++v) ar[v] = ai[v] = static_cast<T>(0);
    for (int32_t e = e0; e < e1; e += G) {
      uint2 en[G];
      int32_t cc[G];
      if constexpr (COOP) {
        const int lane = static_cast<int>(threadIdx.x) & 63, rbase = lane - sub;
        uint2 me = make_uint2(0u, 0u);
        int32_t mc = -1;
        if (sub < G) {
          const int32_t ej = min(e + sub, e1 - 1);
          me = ent[ej];
          const int32_t c = colk[ej];
          mc = (e + sub < e1 && c >= 0) ? c : -1;
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          en[j].x = __shfl(me.x, rbase + j, 64);
          en[j].y = __shfl(me.y, rbase + j, 64);
          cc[j] = __shfl(mc, rbase + j, 64);
        }
      } else {
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int32_t ej = min(e + j, e1 - 1);
          en[j] = ent[ej];
          const int32_t c = colk[ej];
          cc[j] = (e + j < e1 && c >= 0) ? c : -1;
        }
      }
      V2 src[G][VPL];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const V2* sp = x + static_cast<int64_t>(cc[j] >= 0 ? cc[j] : r) * RT + rr;
#pragma unroll
        for (int v = 0; v < VPL; ++v) src[j][v] = sp[v];
      }
#pragma unroll
      for (int j = 0; j < G; ++j) {
        double cr, ci;
        kb_term_coef(en[j], dg, s_nrm, s_coef, s_chi, &cr, &ci);
        if (cc[j] < 0) cr = ci = 0.0;
        const T tr = static_cast<T>(cr), ti = static_cast<T>(ci);
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
          ar[v] = fma(tr, src[j][v].x, ar[v]);
          ai[v] = fma(tr, src[j][v].y, ai[v]);
          ar[v] = fma(-ti, src[j][v].y, ar[v]);
          ai[v] = fma(ti, src[j][v].x, ai[v]);
        }
      }
    }
    V2* dst = y + r * RT + rr;
    const V2* own = x + r * RT + rr;
    const bool add = acc != nullptr && acc[g] != 0;  // the row's down hops are in y already
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
      V2 o;
      o.x = static_cast<T>(static_cast<double>(ar[v]) * sxv[v]);
      o.y = static_cast<T>(static_cast<double>(ai[v]) * sxv[v]);
      if (add) {
        const V2 p = dst[v];
        o.x = p.x + o.x;
        o.y = p.y + o.y;
      }
      dst[v] = o;
      if (dotp) {
        const V2 q = own[v];
        const double qr = sxv[v] * static_cast<double>(q.x), qi = sxv[v] * static_cast<double>(q.y);
        dre[v] += qr * static_cast<double>(o.x) + qi * static_cast<double>(o.y);
        dimg[v] += qr * static_cast<double>(o.y) - qi * static_cast<double>(o.x);
      }
    }
  }
  if (dotp) {  // lane t holds vectors rr .. rr + VPL of its row slot; sum over the L-strided lanes
    __shared__ double red[256][VPL][2];
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
      red[threadIdx.x][v][0] = dre[v];
      red[threadIdx.x][v][1] = dimg[v];
    }
    __syncthreads();
    if (threadIdx.x < RB) {
      const int s2 = threadIdx.x / VPL, v2 = threadIdx.x % VPL;
      double a = 0.0, b = 0.0;
      for (int j = s2; j < 256; j += L) {
        a += red[j][v2][0];
        b += red[j][v2][1];
      }
      dotp[static_cast<int64_t>(blockIdx.x) * RT + r0 + threadIdx.x] = make_double2(a, b);
    }
  }
}


// ------------------------------------------------------------------ momentum stacks
// One apply for every block of a stack.  The R = nk * rk slots of an orbit are spread over L lanes
// of VPL slots (16 bytes, or 8 for an odd count of complex64 slots); a workgroup covers 256 / L
// orbits, grid-strided.  Each term of the K-independent CSR is read once for all blocks: its column
// orbit's membership bits (memc[e] = mem[b], stored per term so that they stream with the entries)
// say in which blocks it is a row, and slot (j, r) takes the
// per-block coefficient ((c_bond * (sign * sqrt(|Stab_row| / |Stab_col|))) * chi*_Kj(R)) computed as
// k_kblock_apply computes it, or 0 off block j.  Terms are added in CSR order, as there, and an
// orbit's slots of the blocks it is not a row of are written as 0.
// ONEJ: every lane's VPL slots belong to one block (rk a multiple of VPL), so a term's coefficient
// is computed once per lane.
// COOP (L a power of two >= 8, so that a row's lanes sit in one wavefront): lanes 0 .. G-1 of a
// row load the entries and membership bits of the row's next G terms with one load each, and every
// lane takes term j's from lane j of its row by a cross-lane read, in place of G loads per lane of
// the same addresses (the vector-memory data path is what these applies saturate).
template <typename T, int VPL, bool ONEJ, bool COOP>
__global__ __launch_bounds__(256) void k_kstack_apply(
    int64_t nrep, int nk, int rk, const uint16_t* __restrict__ mem, const uint16_t* __restrict__ memc,
    const int32_t* __restrict__ row_off, const uint2* __restrict__ ent, const double* __restrict__ diag,
    const double2* __restrict__ coef, int nbond,
    const double2* __restrict__ chi, const uint8_t* __restrict__ acc, const typename Cx<T>::v2* __restrict__ x,
    typename Cx<T>::v2* __restrict__ y, const double* __restrict__ xs, double2* __restrict__ dotp) {
  using V2 = typename Cx<T>::v2;
  constexpr int G = 8;
  const int W = nk * rk;
  const int L = W / VPL;
  const int RPB = 256 / L;
  __shared__ double2 s_coef[kMaxBondsLds];
  __shared__ double2 s_chi[kMaxStackBlocks][kMaxOrbitSites];
  __shared__ double s_nrm[kMaxOrbitSites * kMaxOrbitSites];
  for (int b = threadIdx.x; b < kMaxBondsLds; b += blockDim.x) s_coef[b] = b < nbond ? coef[b] : make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < kMaxStackBlocks * kMaxOrbitSites; i += blockDim.x)
    s_chi[i / kMaxOrbitSites][i % kMaxOrbitSites] = i < nk * kMaxOrbitSites ? chi[i] : make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i < kMaxOrbitSites * kMaxOrbitSites; i += blockDim.x)
    s_nrm[i] = sqrt(static_cast<double>((i & 15) + 1) / static_cast<double>((i >> 4) + 1));
  __syncthreads();
  const int sub = static_cast<int>(threadIdx.x) % L, rl = static_cast<int>(threadIdx.x) / L;
  const int v0 = sub * VPL;
  int jb[VPL];
  double dre[VPL], dimg[VPL], sxv[VPL];
#pragma unroll
  for (int v = 0; v < VPL; ++v) {
    jb[v] = (v0 + v) / rk;
    dre[v] = dimg[v] = 0.0;
    sxv[v] = xs ? xs[v0 + v] : 1.0;
  }
  if (rl < RPB) {
    const RowWalk rw = row_walk(nrep, RPB, rl);
    for (int64_t r = rw.first; r < rw.end; r += rw.stride) {
      const uint32_t mrow = mem[r];
      const int32_t e0 = row_off[r], e1 = row_off[r + 1];
      const double dg = diag[r];
      T ar[VPL], ai[VPL];
#pragma unroll
      for (int v = 0; v < VPL; ++v) ar[v] = ai[v] = static_cast<T>(0);
      for (int32_t e = e0; e < e1; e += G) {
        uint2 en[G];
        uint32_t mb[G];
        if constexpr (COOP) {
          const int lane = static_cast<int>(threadIdx.x) & 63, rbase = lane - sub;
          uint2 me = make_uint2(0u, 0u);
          uint32_t mm = 0u;
          if (sub < G) {
            const int32_t ej = min(e + sub, e1 - 1);
            me = ent[ej];
            mm = e + sub < e1 ? static_cast<uint32_t>(memc[ej]) : 0u;
          }
#pragma unroll
          for (int j = 0; j < G; ++j) {
            en[j].x = __shfl(me.x, rbase + j, 64);
            en[j].y = __shfl(me.y, rbase + j, 64);
            mb[j] = __shfl(mm, rbase + j, 64);
          }
        } else {
#pragma unroll
          for (int j = 0; j < G; ++j) en[j] = ent[min(e + j, e1 - 1)];
#pragma unroll
          for (int j = 0; j < G; ++j) mb[j] = e + j < e1 ? static_cast<uint32_t>(memc[min(e + j, e1 - 1)]) : 0u;
        }
        V2 src[G][VPL];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const V2* sp = x + static_cast<int64_t>(en[j].x & kColMask) * W + v0;
#pragma unroll
          for (int v = 0; v < VPL; ++v) src[j][v] = sp[v];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const bool dgl = (en[j].x & kDiagFlag) != 0;
          const double nrm = s_nrm[(en[j].y >> kStabShift) & 0xFFu];
          const double w = ((en[j].x >> 28) & 1u) ? -nrm : nrm;
          const double2 cb = s_coef[en[j].y & kBondMask];
          const double hr = cb.x * w, hi = cb.y * w;
          const int t = (en[j].x >> 24) & 15u;
          T trv[VPL], tiv[VPL];
#pragma unroll
          for (int v = 0; v < (ONEJ ? 1 : VPL); ++v) {
            const double2 ch = s_chi[jb[v]][t];
            const double hr2 = hr * ch.x - hi * ch.y;
            const double hi2 = hr * ch.y + hi * ch.x;
            double cr = dgl ? dg : hr2, ci = dgl ? 0.0 : hi2;
            if (!((mb[j] >> jb[v]) & 1u)) cr = ci = 0.0;
            trv[v] = static_cast<T>(cr);
            tiv[v] = static_cast<T>(ci);
          }
#pragma unroll
          for (int v = 0; v < VPL; ++v) {
            const T tr = trv[ONEJ ? 0 : v], ti = tiv[ONEJ ? 0 : v];
            ar[v] = fma(tr, src[j][v].x, ar[v]);
            ai[v] = fma(tr, src[j][v].y, ai[v]);
            ar[v] = fma(-ti, src[j][v].y, ar[v]);
            ai[v] = fma(ti, src[j][v].x, ai[v]);
          }
        }
      }
      V2* dst = y + r * W + v0;
      const V2* own = x + r * W + v0;
      const bool add = acc != nullptr && acc[r] != 0;  // the row's down hops are in y already
#pragma unroll
      for (int v = 0; v < VPL; ++v) {
        V2 o;
        if ((mrow >> jb[v]) & 1u) {
          o.x = static_cast<T>(static_cast<double>(ar[v]) * sxv[v]);
          o.y = static_cast<T>(static_cast<double>(ai[v]) * sxv[v]);
          if (add) {
            const V2 p = dst[v];
            o.x = p.x + o.x;
            o.y = p.y + o.y;
          }
        } else {
          o.x = o.y = static_cast<T>(0);
        }
        dst[v] = o;
        if (dotp) {
          const V2 q = own[v];
          const double qr = sxv[v] * static_cast<double>(q.x), qi = sxv[v] * static_cast<double>(q.y);
          dre[v] += qr * static_cast<double>(o.x) + qi * static_cast<double>(o.y);
          dimg[v] += qr * static_cast<double>(o.y) - qi * static_cast<double>(o.x);
        }
      }
    }
  }
  if (dotp) {  // slot v0 + v of every row slot; sum over the L-strided lanes
    __shared__ double red[256][2][2];
#pragma unroll
    for (int v = 0; v < VPL; ++v) {
      red[threadIdx.x][v][0] = rl < RPB ? dre[v] : 0.0;
      red[threadIdx.x][v][1] = rl < RPB ? dimg[v] : 0.0;
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < W) {
      const int s2 = threadIdx.x / VPL, v2 = threadIdx.x % VPL;
      double a = 0.0, b = 0.0;
      for (int j = s2; j < RPB * L; j += L) {
        a += red[j][v2][0];
        b += red[j][v2][1];
      }
      dotp[static_cast<int64_t>(blockIdx.x) * W + threadIdx.x] = make_double2(a, b);
    }
  }
}

// Membership bits of every orbit: bit j set where the orbit is a row of block j.
__global__ void k_kstack_mem(int64_t nrep, int nk, KStackMaps maps, uint16_t* mem) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nrep) return;
  uint32_t m = 0;
  for (int j = 0; j < nk; ++j) m |= (maps.map[j][g] >= 0 ? 1u : 0u) << j;
  mem[g] = static_cast<uint16_t>(m);
}

// Membership bits of every term's column orbit, in term order.
__global__ void k_kstack_memc(const uint2* ent, int64_t nnz, const uint16_t* mem, uint16_t* memc) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < nnz) memc[e] = mem[ent[e].x & kColMask];
}

// Block j's vectors [dim_j][rk] into / out of slots (j, 0 .. rk) of a stacked block [nrep][W].
// Scatter writes every orbit's slots (0 off the block).
template <typename T>
__global__ void k_kstack_scatter(int64_t nrep, int W, int j, int rk, const int32_t* __restrict__ map,
                                 const typename Cx<T>::v2* __restrict__ xb, typename Cx<T>::v2* __restrict__ xs) {
  using V2 = typename Cx<T>::v2;
  const int64_t n = nrep * rk;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t g = f / rk;
    const int r = static_cast<int>(f - g * rk);
    const int32_t i = map[g];
    V2 v;
    if (i >= 0) {
      v = xb[static_cast<int64_t>(i) * rk + r];
    } else {
      v.x = v.y = static_cast<T>(0);
    }
    xs[g * W + static_cast<int64_t>(j) * rk + r] = v;
  }
}

template <typename T>
__global__ void k_kstack_gather(int64_t nrep, int W, int j, int rk, const int32_t* __restrict__ map,
                                const typename Cx<T>::v2* __restrict__ xs, typename Cx<T>::v2* __restrict__ xb) {
  const int64_t n = nrep * rk;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t g = f / rk;
    const int r = static_cast<int>(f - g * rk);
    const int32_t i = map[g];
    if (i >= 0) xb[static_cast<int64_t>(i) * rk + r] = xs[g * W + static_cast<int64_t>(j) * rk + r];
  }
}


// Term-split bilinear forms of a momentum block (derivative operators dH_p over the same block):
// per hop table j (derivative coefficients over this graph's bond ids, [j][kMaxBondsLds])
// Re <a| dH_j |b>, and per diagonal array k (a derivative operator's own orbit diagonal)
// Re <a| diag_k |b>, summed over the R vectors.  Terms are the apply's: coefficient x sign x
// sqrt(|Stab_row| / |Stab_col|) x Bloch phase, the column resolved per K.  One thread per
// (row, vector).  partials: part[block * kKbQ + j] (hop tables) and [block * kKbQ + 4 + k].
constexpr int kKbTables = 4;
constexpr int kKbQ = 2 * kKbTables;

struct KbDiags {
  const double* d[kKbTables];
};

__global__ __launch_bounds__(256) void k_bil_kblock(const int32_t* __restrict__ rows, int64_t dk,
                                                    const int32_t* __restrict__ row_off,
                                                    const uint2* __restrict__ ent, const int32_t* __restrict__ colk,
                                                    Phases ph, const double2* __restrict__ dc, int ntab, KbDiags dg,
                                                    int ndg, const double2* __restrict__ a,
                                                    const double2* __restrict__ b, int R, double2* __restrict__ part) {
  __shared__ double2 s_dc[kKbTables][kMaxBondsLds];
  __shared__ double2 s_chi[kMaxOrbitSites];
  __shared__ double s_nrm[kMaxOrbitSites * kMaxOrbitSites];
  for (int i = threadIdx.x; i < kKbTables * kMaxBondsLds; i += blockDim.x)
    s_dc[i / kMaxBondsLds][i % kMaxBondsLds] = i / kMaxBondsLds < ntab ? dc[i] : make_double2(0.0, 0.0);
  if (threadIdx.x < kMaxOrbitSites) s_chi[threadIdx.x] = ph.chi[threadIdx.x];
  for (int i = threadIdx.x; i < kMaxOrbitSites * kMaxOrbitSites; i += blockDim.x)
    s_nrm[i] = sqrt(static_cast<double>((i & 15) + 1) / static_cast<double>((i >> 4) + 1));
  __syncthreads();
  double acc[kKbQ];
#pragma unroll
  for (int j = 0; j < kKbQ; ++j) acc[j] = 0.0;
  const int64_t n = dk * R;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t r = f / R;
    const int v = static_cast<int>(f - r * R);
    const int32_t g = rows[r];
    const double2 av = a[f];
    double tr[kKbTables], ti[kKbTables];
#pragma unroll
    for (int j = 0; j < kKbTables; ++j) tr[j] = ti[j] = 0.0;
    double dre = 0.0;  // Re conj(a) b over the diagonal entry
    const int32_t e0 = row_off[g], e1 = row_off[g + 1];
    constexpr int G = 8;  // terms whose loads are issued together
    for (int32_t e = e0; e < e1; e += G) {
      uint2 en[G];
      int32_t cc[G];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const int32_t ej = min(e + j, e1 - 1);
        en[j] = ent[ej];
        const int32_t c = colk[ej];
        cc[j] = (e + j < e1 && c >= 0) ? c : -1;
      }
      double2 src[G];
#pragma unroll
      for (int j = 0; j < G; ++j) src[j] = b[(cc[j] >= 0 ? static_cast<int64_t>(cc[j]) : r) * R + v];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        if (cc[j] < 0) continue;
        if (en[j].x & kDiagFlag) {
          dre += av.x * src[j].x + av.y * src[j].y;
          continue;
        }
        const double nrm = s_nrm[(en[j].y >> kStabShift) & 0xFFu];
        const double w = ((en[j].x >> 28) & 1u) ? -nrm : nrm;
        const double2 ch = s_chi[(en[j].x >> 24) & 15u];
        const uint32_t bd = en[j].y & kBondMask;
#pragma unroll
        for (int t = 0; t < kKbTables; ++t)
          if (t < ntab) {
            const double2 cb = s_dc[t][bd];
            const double hr = cb.x * w, hi = cb.y * w;
            const double cr = hr * ch.x - hi * ch.y, ci = hr * ch.y + hi * ch.x;
            tr[t] += cr * src[j].x - ci * src[j].y;
            ti[t] += cr * src[j].y + ci * src[j].x;
          }
      }
    }
#pragma unroll
    for (int j = 0; j < kKbTables; ++j) acc[j] += av.x * tr[j] + av.y * ti[j];
#pragma unroll
    for (int k = 0; k < kKbTables; ++k)
      if (k < ndg) acc[kKbTables + k] += dg.d[k][g] * dre;
  }
  __shared__ double red[256 / 64][kKbQ];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kKbQ; ++j) {
    double x = acc[j];
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[wv][j] = x;
  }
  __syncthreads();
  if (threadIdx.x < kKbQ) {
    double x = 0.0;
    for (int w = 0; w < 256 / 64; ++w) x += red[w][threadIdx.x];
    part[static_cast<int64_t>(blockIdx.x) * kKbQ + threadIdx.x] = make_double2(x, 0.0);
  }
}

// ------------------------------------------------------------------ host objects
namespace {

std::vector<uint32_t> strings_numeric(int n, int k) {
  std::vector<uint32_t> v;
  if (k == 0) {
    v.push_back(0);
    return v;
  }
  if (k > n) return v;
  uint32_t m = (k >= 32) ? 0xFFFFFFFFu : ((1u << k) - 1u);
  const uint64_t lim = uint64_t{1} << n;
  while (m < lim) {
    v.push_back(m);
    const uint32_t c = m & (0u - m), r = m + c;  // next mask with the same popcount (Gosper)
    if (r == 0) break;
    m = (((r ^ m) >> 2) / c) | r;
  }
  return v;
}

template <class F>
void cub_call(F&& f, hipStream_t st) {
  size_t bytes = 0;
  FTLM_HIP_CHECK(f(nullptr, bytes, st));
  DevBuf<uint8_t> tmp(std::max<size_t>(bytes, 1));
  FTLM_HIP_CHECK(f(tmp.p, bytes, st));
  FTLM_HIP_CHECK(hipStreamSynchronize(st));
}

// Block row of each trivial line's first orbit (the line's orbits are rows, consecutive).
__global__ void k_line_rows(const int64_t* lbase, int nlines, const int32_t* map, int64_t* out) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < nlines) out[l] = map[lbase[l]];
}

// Terms a split apply leaves to the momentum kernels: all terms of orbits off the trivial lines,
// and the diagonal and up-hop terms of those on them (their down hops go to the sector row pass).  keep[e] per term, cnt[g] = kept terms of row g.
__global__ void k_split_keep(const int32_t* row_off, const uint2* ent, int64_t nrep, const uint8_t* triv,
                             const uint32_t* dnbits, uint8_t* keep, int32_t* cnt) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nrep) return;
  const int32_t e0 = row_off[g], e1 = row_off[g + 1];
  int32_t c = 0;
  for (int32_t e = e0; e < e1; ++e) {
    const uint2 en = ent[e];
    const uint32_t b = en.y & kBondMask;
    const bool down = (en.x & kDiagFlag) == 0 && ((dnbits[(b >> 5) & 15u] >> (b & 31u)) & 1u) != 0;
    const bool k = !(triv[g] && down);
    keep[e] = k ? 1 : 0;
    c += k ? 1 : 0;
  }
  cnt[g] = c;
}

}  // namespace

// Split applies of an orbit graph (orbit_split_build).  In the pack order an orbit's
// representative (u, d) has the smallest up string of its orbit.  Where no translation fixes u, every
// down string d occurs with it: those orbits form a trivial line of D_dn consecutive orbits in
// increasing-mask order of d, and a down hop inside the line is the full sector's row-u down hop
// (identity translation, unit norm, the same Jordan-Wigner signs, no Bloch phase), in every block,
// of which every orbit of the line is a row.  The lines' down hops go through the full sector's
// row pass (H*V kernels over a sector whose strings are in increasing-mask order); the momentum
// kernels add the remaining terms, kept in a second CSR.
struct OrbitSplit {
  std::unique_ptr<Sector> sec;  // the full sector, strings in increasing-mask order
  int nlines = 0;
  DevBuf<int64_t> lbase;        // [nlines] first orbit of each trivial line
  DevBuf<uint32_t> lumask;      // [nlines] its up string
  DevBuf<uint8_t> triv;         // [nrep] 1 on a trivial line
  DevBuf<int32_t> row_off2;     // [nrep + 1] the remaining terms, in CSR order
  DevBuf<uint2> ent2;
  int64_t nnz2 = 0;
};

struct OrbitGraph {
  int lx = 0, ly = 0, n_up = 0, n_dn = 0;
  double U = 0.0, V = 0.0;
  std::vector<ftlm_hop_t> hops;
  std::vector<ftlm_pair_t> pairs;
  LatDev L{};
  int64_t nrep = 0, nnz = 0;
  int nbond = 0;
  bool real_h = true;
  bool hermitian = true;  // the merged hop coefficients are translation invariant
  DevBuf<uint32_t> pack;
  DevBuf<uint16_t> stab, eta;
  DevBuf<double> diag;
  DevBuf<int32_t> row_off;
  DevBuf<uint2> ent;
  DevBuf<double2> coef;
  std::unique_ptr<OrbitSplit> split;  // null: no split applies (see OrbitSplit)

  bool same_model(int lx_, int ly_, int nu, int nd, double U_, double V_, const ftlm_hop_t* h, int nh,
                  const ftlm_pair_t* p, int np) const {
    if (lx_ != lx || ly_ != ly || nu != n_up || nd != n_dn || U_ != U || V_ != V) return false;
    if (nh != static_cast<int>(hops.size()) || np != static_cast<int>(pairs.size())) return false;
    for (int i = 0; i < nh; ++i)
      if (h[i].from != hops[i].from || h[i].to != hops[i].to || h[i].spin != hops[i].spin || h[i].re != hops[i].re ||
          h[i].im != hops[i].im)
        return false;
    for (int i = 0; i < np; ++i)
      if (p[i].i != pairs[i].i || p[i].j != pairs[i].j) return false;
    return true;
  }
};

struct KBlock : ftlm_op {
  std::shared_ptr<OrbitGraph> G;
  int kx = 0, ky = 0;
  double tol = 1e-12;
  Phases ph{};
  DevBuf<int32_t> rows, map;
  DevBuf<int32_t> colk;  // [nnz] block column of every stored term, -1 when not a row here
  DevBuf<int32_t> colk2;  // [nnz2] the same for the split's remaining terms
  DevBuf<int64_t> lbase;  // [nlines] block row of the split's trivial lines' first orbits
  DevBuf<double> sq, isq;
  DevBuf<double2> A;  // [nrep]
};

// True when the merged hop coefficients (equal (from, to, spin) summed) are invariant under the
// lattice translations (site = y * lx + x); the block operator is Hermitian exactly then.  The
// reference's Peierls phases sit on the wrapped bonds only, so a nonzero flux breaks this: its
// block operator then has |H - H^dag| of the order of the flux and complex eigenvalues.
bool hops_translation_invariant(int lx, int ly, const ftlm_hop_t* hops, int n_hops) {
  std::map<std::tuple<int, int, int>, std::complex<double>> m;
  double scale = 0.0;
  for (int h = 0; h < n_hops; ++h) {
    if (hops[h].from == hops[h].to) continue;
    m[std::make_tuple(hops[h].from, hops[h].to, hops[h].spin)] += std::complex<double>(hops[h].re, hops[h].im);
  }
  for (const auto& kv : m) scale = std::max(scale, std::abs(kv.second));
  const double tol = 1e-12 * std::max(scale, 1.0);
  auto tr = [&](int site, int dx, int dy) { return ((site / lx + dy) % ly) * lx + (site % lx + dx) % lx; };
  const int gen[2][2] = {{1, 0}, {0, 1}};
  for (const auto& g : gen)
    for (const auto& kv : m) {
      const auto key = std::make_tuple(tr(std::get<0>(kv.first), g[0], g[1]), tr(std::get<1>(kv.first), g[0], g[1]),
                                       std::get<2>(kv.first));
      const auto it = m.find(key);
      const std::complex<double> c = it == m.end() ? std::complex<double>(0.0, 0.0) : it->second;
      if (std::abs(c - kv.second) > tol) return false;
    }
  return true;
}

// Whether orbit graphs get split applies (FTLM_ORBIT_SPLIT=0 keeps the one-kernel applies).
bool orbit_split_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("FTLM_ORBIT_SPLIT");
    return v == nullptr || std::atoi(v) != 0;
  }();
  return on;
}

// Builds G->split (see OrbitSplit) when the bond coefficients are real and the full sector has
// D_up, D_dn >= 128 and some trivial line; otherwise leaves it null.
void orbit_split_build(OrbitGraph* Gp) {
  OrbitGraph& G = *Gp;
  if (!orbit_split_enabled() || !G.real_h || G.nrep == 0 || G.nnz == 0) return;
  const int n = G.lx * G.ly;
  if (binom(n, G.n_up) < 128 || binom(n, G.n_dn) < 128) return;
  auto sp = std::make_unique<OrbitSplit>();
  sp->sec = sector_create_numeric(n, G.n_up, G.n_dn, G.hops.data(), static_cast<int>(G.hops.size()), G.pairs.data(),
                                  static_cast<int>(G.pairs.size()), G.U, G.V);
  const Sector* Sx = sp->sec.get();
  if (!Sx->real_h || Sx->Ddn < 128) return;
  // trivial lines, from the representatives (pack order: up string, then down string)
  const int64_t nr = G.nrep;
  std::vector<uint32_t> pack(static_cast<size_t>(nr));
  FTLM_HIP_CHECK(hipMemcpy(pack.data(), G.pack.p, sizeof(uint32_t) * pack.size(), hipMemcpyDeviceToHost));
  std::vector<uint32_t> dn(static_cast<size_t>(Sx->Ddn));
  FTLM_HIP_CHECK(hipMemcpy(dn.data(), Sx->dn_mask.p, sizeof(uint32_t) * dn.size(), hipMemcpyDeviceToHost));
  std::vector<int64_t> lbase;
  std::vector<uint32_t> lumask;
  std::vector<uint8_t> triv(static_cast<size_t>(nr), 0);
  for (int64_t a = 0; a < nr;) {
    const uint32_t u = pack[static_cast<size_t>(a)] >> 16;
    int64_t b = a;
    while (b < nr && (pack[static_cast<size_t>(b)] >> 16) == u) ++b;
    bool full = b - a == Sx->Ddn;
    for (int64_t i = a; full && i < b; ++i)
      full = (pack[static_cast<size_t>(i)] & 0xFFFFu) == dn[static_cast<size_t>(i - a)];
    if (full) {
      lbase.push_back(a);
      lumask.push_back(u);
      std::fill(triv.begin() + a, triv.begin() + b, 1);
    }
    a = b;
  }
  if (lbase.empty()) return;
  // down-spin bond ids: the merged bond list (first occurrence of (from, to, spin), self hops
  // skipped), as build_orbit_graph numbers them
  std::vector<uint32_t> dnbits(16, 0u);
  {
    std::map<std::tuple<int, int, int>, int> seen;
    int nb = 0;
    for (const ftlm_hop_t& e : G.hops) {
      if (e.from == e.to) continue;
      const auto key = std::make_tuple(e.from, e.to, e.spin);
      if (seen.count(key)) continue;
      seen.emplace(key, nb);
      if (e.spin == 1) dnbits[static_cast<size_t>(nb >> 5)] |= 1u << (nb & 31);
      ++nb;
    }
    FTLM_REQUIRE(nb == G.nbond, FTLM_EHIP, "orbit graph: bond list mismatch");
  }
  sp->nlines = static_cast<int>(lbase.size());
  sp->lbase.alloc(lbase.size());
  sp->lbase.upload(lbase);
  sp->lumask.alloc(lumask.size());
  sp->lumask.upload(lumask);
  sp->triv.alloc(triv.size());
  sp->triv.upload(triv);
  DevBuf<uint32_t> ddn(16);
  ddn.upload(dnbits);
  DevBuf<uint8_t> keep(static_cast<size_t>(G.nnz));
  DevBuf<int32_t> cnt(static_cast<size_t>(nr));
  const hipStream_t st = nullptr;
  hipLaunchKernelGGL(k_split_keep, dim3(grid_for(nr, 256)), dim3(256), 0, st, G.row_off.p, G.ent.p, nr, sp->triv.p,
                     ddn.p, keep.p, cnt.p);
  check_launch("k_split_keep");
  std::vector<int32_t> c(static_cast<size_t>(nr)), off(static_cast<size_t>(nr) + 1);
  FTLM_HIP_CHECK(hipMemcpy(c.data(), cnt.p, sizeof(int32_t) * c.size(), hipMemcpyDeviceToHost));
  off[0] = 0;
  for (int64_t g = 0; g < nr; ++g)
    off[static_cast<size_t>(g) + 1] = off[static_cast<size_t>(g)] + c[static_cast<size_t>(g)];
  sp->nnz2 = off[static_cast<size_t>(nr)];
  sp->row_off2.alloc(off.size());
  sp->row_off2.upload(off);
  sp->ent2.alloc(static_cast<size_t>(std::max<int64_t>(sp->nnz2, 1)));
  DevBuf<int32_t> nsel(1);
  const int nn = static_cast<int>(G.nnz);
  cub_call([&](void* t, size_t& b, hipStream_t s2) {
    return hipcub::DeviceSelect::Flagged(t, b, G.ent.p, keep.p, sp->ent2.p, nsel.p, nn, s2);
  }, st);
  int32_t got = 0;
  FTLM_HIP_CHECK(hipMemcpy(&got, nsel.p, sizeof(int32_t), hipMemcpyDeviceToHost));
  FTLM_REQUIRE(got == sp->nnz2, FTLM_EHIP, "orbit graph: split term count mismatch");
  G.split = std::move(sp);
}

std::shared_ptr<OrbitGraph> build_orbit_graph(int lx, int ly, int n_up, int n_dn, const ftlm_hop_t* hops, int n_hops,
                                              const ftlm_pair_t* pairs, int n_pairs, double U, double V) {
  const int n = lx * ly;
  auto G = std::make_shared<OrbitGraph>();
  G->lx = lx;
  G->ly = ly;
  G->n_up = n_up;
  G->n_dn = n_dn;
  G->U = U;
  G->V = V;
  G->hops.assign(hops, hops + n_hops);
  G->pairs.assign(pairs, pairs + n_pairs);
  G->L = make_lat(lx, ly);
  G->hermitian = hops_translation_invariant(lx, ly, hops, n_hops);
  const hipStream_t st = nullptr;

  // merged bonds (same (from, to, spin) summed in list order), in first-occurrence order
  std::vector<OrbBond> bonds;
  std::vector<double2> cf;
  std::map<std::tuple<int, int, int>, int> seen;
  for (int h = 0; h < n_hops; ++h) {
    const ftlm_hop_t& e = hops[h];
    FTLM_REQUIRE(e.spin == 0 || e.spin == 1, FTLM_EINVAL, "hop spin must be 0 or 1");
    FTLM_REQUIRE(e.from >= 0 && e.from < n && e.to >= 0 && e.to < n, FTLM_EINVAL, "hop site outside [0, n_sites)");
    if (e.from == e.to) continue;
    const auto key = std::make_tuple(e.from, e.to, e.spin);
    auto it = seen.find(key);
    if (it == seen.end()) {
      seen.emplace(key, static_cast<int>(bonds.size()));
      bonds.push_back(OrbBond{e.from, e.to, e.spin});
      cf.push_back(make_double2(e.re, e.im));
    } else {
      cf[it->second].x += e.re;
      cf[it->second].y += e.im;
    }
  }
  FTLM_REQUIRE(static_cast<int>(bonds.size()) <= kMaxBondsLds, FTLM_EUNSUPPORTED,
               "more than 512 distinct directed bonds");
  for (const double2& c : cf)
    if (c.y != 0.0) G->real_h = false;
  G->nbond = static_cast<int>(bonds.size());
  G->coef.alloc(std::max<size_t>(cf.size(), 1));
  if (!cf.empty()) G->coef.upload(cf);
  DevBuf<OrbBond> dbonds(std::max<size_t>(bonds.size(), 1));
  if (!bonds.empty()) dbonds.upload(bonds.data(), bonds.size());
  std::vector<int2> pr(static_cast<size_t>(std::max(n_pairs, 1)));
  for (int p = 0; p < n_pairs; ++p) {
    FTLM_REQUIRE(pairs[p].i >= 0 && pairs[p].i < n && pairs[p].j >= 0 && pairs[p].j < n, FTLM_EINVAL,
                 "pair site outside [0, n_sites)");
    pr[p] = make_int2(pairs[p].i, pairs[p].j);
  }
  DevBuf<int2> dpairs(pr.size());
  dpairs.upload(pr);

  // 1. representatives, in pack order
  const std::vector<uint32_t> upN = strings_numeric(n, n_up), dnN = strings_numeric(n, n_dn);
  const int Du = static_cast<int>(upN.size()), Dd = static_cast<int>(dnN.size());
  const int64_t dim = static_cast<int64_t>(Du) * Dd;
  FTLM_REQUIRE(dim < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "sector too large for the orbit scan");
  DevBuf<uint32_t> dup(upN.size()), ddn(dnN.size());
  dup.upload(upN);
  ddn.upload(dnN);
  DevBuf<uint8_t> flag(static_cast<size_t>(dim));
  hipLaunchKernelGGL(k_orb_flag, dim3(grid_for(dim, 256)), dim3(256), 0, st, dup.p, Du, ddn.p, Dd, G->L, flag.p);
  check_launch("k_orb_flag");
  DevBuf<int32_t> idx(static_cast<size_t>(dim));
  DevBuf<int32_t> nsel(1);
  {
    hipcub::CountingInputIterator<int32_t> it(0);
    const int ndim = static_cast<int>(dim);
    cub_call([&](void* t, size_t& b, hipStream_t s) {
      return hipcub::DeviceSelect::Flagged(t, b, it, flag.p, idx.p, nsel.p, ndim, s);
    }, st);
  }
  int32_t nr = 0;
  FTLM_HIP_CHECK(hipMemcpy(&nr, nsel.p, sizeof(int32_t), hipMemcpyDeviceToHost));
  flag.release();
  G->nrep = nr;
  FTLM_REQUIRE(nr < static_cast<int32_t>(kColMask), FTLM_EUNSUPPORTED, "more than 2^24 orbits");
  G->pack.alloc(std::max<int64_t>(nr, 1));
  G->stab.alloc(std::max<int64_t>(nr, 1));
  G->eta.alloc(std::max<int64_t>(nr, 1));
  G->diag.alloc(std::max<int64_t>(nr, 1));
  if (nr > 0) {
    hipLaunchKernelGGL(k_orb_reps, dim3(grid_for(nr, 256)), dim3(256), 0, st, idx.p, static_cast<int64_t>(nr),
                       dup.p, ddn.p, Dd, G->L, dpairs.p, n_pairs, U, V, G->pack.p, G->stab.p, G->eta.p, G->diag.p);
    check_launch("k_orb_reps");
  }
  idx.release();

  // 2. push entries, stably sorted by target orbit
  DevBuf<int32_t> cnt(std::max<int64_t>(nr, 1)), off(static_cast<size_t>(nr) + 1);
  if (nr > 0) {
    hipLaunchKernelGGL(k_orb_count, dim3(grid_for(nr, 256)), dim3(256), 0, st, G->pack.p, static_cast<int64_t>(nr),
                       dbonds.p, G->nbond, cnt.p);
    check_launch("k_orb_count");
  }
  std::vector<int32_t> hc(static_cast<size_t>(nr));
  if (nr > 0) FTLM_HIP_CHECK(hipMemcpy(hc.data(), cnt.p, sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
  std::vector<int32_t> ho(static_cast<size_t>(nr) + 1, 0);
  int64_t acc = 0;
  for (int32_t g = 0; g < nr; ++g) {
    ho[g] = static_cast<int32_t>(acc);
    acc += hc[g];
    FTLM_REQUIRE(acc < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "orbit hop table too large");
  }
  ho[nr] = static_cast<int32_t>(acc);
  G->nnz = acc;
  off.upload(ho);
  DevBuf<uint32_t> k0(std::max<int64_t>(acc, 1)), k1(std::max<int64_t>(acc, 1));
  DevBuf<uint2> v0(std::max<int64_t>(acc, 1)), v1(std::max<int64_t>(acc, 1));
  DevBuf<int> bad(1);
  FTLM_HIP_CHECK(hipMemset(bad.p, 0, sizeof(int)));
  if (nr > 0) {
    hipLaunchKernelGGL(k_orb_fill, dim3(grid_for(nr, 256)), dim3(256), 0, st, G->pack.p, G->stab.p,
                       static_cast<int64_t>(nr), dbonds.p, G->nbond, off.p, G->L, k0.p, v0.p, bad.p);
    check_launch("k_orb_fill");
  }
  int hbad = 0;
  FTLM_HIP_CHECK(hipMemcpy(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost));
  FTLM_REQUIRE(hbad == 0, FTLM_EHIP, "orbit graph: a hop left the orbit list");
  int end_bit = 1;
  while ((int64_t{1} << end_bit) < std::max<int64_t>(nr, 2)) ++end_bit;
  hipcub::DoubleBuffer<uint32_t> K(k0.p, k1.p);
  hipcub::DoubleBuffer<uint2> Vv(v0.p, v1.p);
  if (acc > 0) {
    const int nn = static_cast<int>(acc);
    cub_call([&](void* t, size_t& b, hipStream_t s) {
      return hipcub::DeviceRadixSort::SortPairs(t, b, K, Vv, nn, 0, end_bit, s);
    }, st);
  }
  G->row_off.alloc(static_cast<size_t>(nr) + 1);
  if (acc > 0) {
    hipLaunchKernelGGL(k_csr_rows, dim3(grid_for(acc, 256)), dim3(256), 0, st, K.Current(), acc, G->row_off.p);
    check_launch("k_csr_rows");
  }
  const int32_t last = static_cast<int32_t>(acc);
  FTLM_HIP_CHECK(hipMemcpy(G->row_off.p + nr, &last, sizeof(int32_t), hipMemcpyHostToDevice));
  // keep the sorted values; release the rest
  if (Vv.Current() == v0.p) {
    G->ent = std::move(v0);
  } else {
    G->ent = std::move(v1);
  }
  FTLM_HIP_CHECK(hipDeviceSynchronize());
  orbit_split_build(G.get());
  return G;
}

void build_kblock(KBlock* B, double tol) {
  const OrbitGraph& G = *B->G;
  const int64_t nr = G.nrep;
  const hipStream_t st = nullptr;
  // translation_bloch_phase (include/ftlm/symmetry/momentum_sector.hpp:57-63), host libm
  const double pi = std::acos(-1.0);
  for (int dy = 0; dy < G.ly; ++dy)
    for (int dx = 0; dx < G.lx; ++dx) {
      const double th = -2.0 * pi *
                        (static_cast<double>(B->kx * dx) / static_cast<double>(G.lx) +
                         static_cast<double>(B->ky * dy) / static_cast<double>(G.ly));
      B->ph.chi[dy * G.lx + dx] = make_double2(std::cos(th), std::sin(th));
    }
  B->A.alloc(std::max<int64_t>(nr, 1));
  B->map.alloc(std::max<int64_t>(nr, 1));
  B->sq.alloc(std::max<int64_t>(nr, 1));
  B->isq.alloc(std::max<int64_t>(nr, 1));
  DevBuf<uint8_t> member(std::max<int64_t>(nr, 1));
  DevBuf<int32_t> rows(std::max<int64_t>(nr, 1)), nsel(1);
  int32_t dk = 0;
  if (nr > 0) {
    hipLaunchKernelGGL(k_kblock_A, dim3(grid_for(nr, 256)), dim3(256), 0, st, G.stab.p, G.eta.p, nr, G.L.nt, B->ph,
                       tol, B->A.p, member.p);
    check_launch("k_kblock_A");
    hipcub::CountingInputIterator<int32_t> it(0);
    const int nn = static_cast<int>(nr);
    cub_call([&](void* t, size_t& b, hipStream_t s) {
      return hipcub::DeviceSelect::Flagged(t, b, it, member.p, rows.p, nsel.p, nn, s);
    }, st);
    FTLM_HIP_CHECK(hipMemcpy(&dk, nsel.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nr, 256)), dim3(256), 0, st, B->map.p, nr, -1);
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(nr, 256)), dim3(256), 0, st, B->sq.p, nr, 0.0);
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(nr, 256)), dim3(256), 0, st, B->isq.p, nr, 0.0);
    check_launch("k_fill");
  }
  B->rows.alloc(std::max<int32_t>(dk, 1));
  if (dk > 0) {
    FTLM_HIP_CHECK(hipMemcpy(B->rows.p, rows.p, sizeof(int32_t) * dk, hipMemcpyDeviceToDevice));
    hipLaunchKernelGGL(k_kblock_maps, dim3(grid_for(dk, 256)), dim3(256), 0, st, B->rows.p,
                       static_cast<int64_t>(dk), B->A.p, B->map.p, B->sq.p, B->isq.p);
    check_launch("k_kblock_maps");
    B->colk.alloc(static_cast<size_t>(std::max<int64_t>(G.nnz, 1)));
    if (G.nnz > 0) {
      hipLaunchKernelGGL(k_kblock_cols, dim3(grid_for(G.nnz, 256)), dim3(256), 0, st, G.ent.p, G.nnz, B->map.p,
                         B->colk.p);
      check_launch("k_kblock_cols");
    }
    if (G.split) {  // every orbit of a trivial line is a row of every block, in one run of rows
      const OrbitSplit& sp = *G.split;
      B->colk2.alloc(static_cast<size_t>(std::max<int64_t>(sp.nnz2, 1)));
      if (sp.nnz2 > 0) {
        hipLaunchKernelGGL(k_kblock_cols, dim3(grid_for(sp.nnz2, 256)), dim3(256), 0, st, sp.ent2.p, sp.nnz2,
                           B->map.p, B->colk2.p);
        check_launch("k_kblock_cols");
      }
      B->lbase.alloc(static_cast<size_t>(sp.nlines));
      hipLaunchKernelGGL(k_line_rows, dim3(grid_for(sp.nlines, 256)), dim3(256), 0, st, sp.lbase.p, sp.nlines,
                         B->map.p, B->lbase.p);
      check_launch("k_line_rows");
    }
  }
  FTLM_HIP_CHECK(hipDeviceSynchronize());
  B->dim = dk;
}

namespace {

template <typename T, int RB>
void launch_kblock(const KBlock* B, const void* x, void* y, int RT, int r0, const double* xs, double2* dotp,
                   hipStream_t st, int grid) {
  using V2 = typename Cx<T>::v2;
  const OrbitGraph& G = *B->G;
  const OrbitSplit* sp = G.split.get();  // with a split, the remaining terms (added on the lines)
  const int32_t* roff = sp ? sp->row_off2.p : G.row_off.p;
  const uint2* ents = sp ? sp->ent2.p : G.ent.p;
  const int32_t* cols = sp ? B->colk2.p : B->colk.p;
  const uint8_t* acc = sp ? sp->triv.p : nullptr;
  if (RB * sizeof(V2) >= 128) {  // 8 or more lanes per row: shared entry loads
    hipLaunchKernelGGL((k_kblock_apply_split<T, RB, true>), dim3(grid), dim3(256), 0, st, B->rows.p, B->dim, roff,
                       ents, cols, G.diag.p, G.coef.p, G.nbond, B->ph, acc, static_cast<const V2*>(x),
                       static_cast<V2*>(y), RT, r0, xs, dotp);
  } else if (RB * sizeof(V2) >= 32) {
    hipLaunchKernelGGL((k_kblock_apply_split<T, RB, false>), dim3(grid), dim3(256), 0, st, B->rows.p, B->dim, roff,
                       ents, cols, G.diag.p, G.coef.p, G.nbond, B->ph, acc, static_cast<const V2*>(x),
                       static_cast<V2*>(y), RT, r0, xs, dotp);
  } else {
    hipLaunchKernelGGL((k_kblock_apply<T, RB>), dim3(grid), dim3(256), 0, st, B->rows.p, B->dim, roff, ents, cols,
                       G.diag.p, G.coef.p, G.nbond, B->ph, acc, static_cast<const V2*>(x), static_cast<V2*>(y), RT,
                       r0, xs, dotp);
  }
}

template <typename T>
void launch_kblock_chunk(const KBlock* B, const void* x, void* y, int RT, int r0, int rb, const double* xs,
                         double2* dotp, hipStream_t st, int grid) {
  switch (rb) {
    case 16: launch_kblock<T, 16>(B, x, y, RT, r0, xs, dotp, st, grid); break;
    case 8: launch_kblock<T, 8>(B, x, y, RT, r0, xs, dotp, st, grid); break;
    case 4: launch_kblock<T, 4>(B, x, y, RT, r0, xs, dotp, st, grid); break;
    case 2: launch_kblock<T, 2>(B, x, y, RT, r0, xs, dotp, st, grid); break;
    default: launch_kblock<T, 1>(B, x, y, RT, r0, xs, dotp, st, grid); break;
  }
}

}  // namespace

constexpr int kKbMaxGrid = 4096;
// Momentum-block applies walk their rows grid-strided, so the rows in flight span grid * 256
// consecutive orbits; grids of 256-512 workgroups beat 4096 by 8-35% at 4x4 (8,8) for R = 1..32
// (timed by the autotuner), the smaller window likely keeping more shared sources in L2.
constexpr int kKbDefaultGrid = 512;

bool kblock_hermitian(const ftlm_op* op) { return static_cast<const KBlock*>(op)->G->hermitian; }

bool kblock_termsplit(const ftlm_op* op, ftlm_op* const* dops, int npar, KbTermSplit* T) {
  if (op->kind != kKBlock) return false;
  const KBlock* H = static_cast<const KBlock*>(op);
  const OrbitGraph& G = *H->G;
  // H's bond ids: merged (from, to, spin) in first-occurrence order, as build_orbit_graph numbers them
  std::map<std::tuple<int, int, int>, int> id;
  for (const ftlm_hop_t& h : G.hops) {
    if (h.from == h.to) continue;
    const auto key = std::make_tuple(h.from, h.to, h.spin);
    if (id.find(key) == id.end()) {
      const int b = static_cast<int>(id.size());
      id.emplace(key, b);
    }
  }
  std::vector<std::complex<double>> tabs;
  T->tab.assign(static_cast<size_t>(npar), -1);
  T->dg.assign(static_cast<size_t>(npar), -1);
  T->ntab = T->ndg = 0;
  for (int p = 0; p < npar; ++p) {
    if (dops[p]->kind != kKBlock) return false;
    const KBlock* D = static_cast<const KBlock*>(dops[p]);
    const OrbitGraph& GD = *D->G;
    if (GD.lx != G.lx || GD.ly != G.ly || GD.n_up != G.n_up || GD.n_dn != G.n_dn || D->kx != H->kx ||
        D->ky != H->ky || D->dim != H->dim || GD.nrep != G.nrep)
      return false;
    bool any_hop = false;
    for (const ftlm_hop_t& h : GD.hops)
      if (h.from != h.to) any_hop = true;
    if (any_hop) {
      if (T->ntab == kKbTables) return false;
      const int j = T->ntab++;
      tabs.resize(static_cast<size_t>(T->ntab) * kMaxBondsLds);
      for (const ftlm_hop_t& h : GD.hops) {
        if (h.from == h.to) continue;
        const auto it = id.find(std::make_tuple(h.from, h.to, h.spin));
        if (it == id.end()) return false;
        tabs[static_cast<size_t>(j) * kMaxBondsLds + it->second] += std::complex<double>(h.re, h.im);
      }
      T->tab[p] = j;
    }
    if (GD.U != 0.0 || GD.V != 0.0) {
      if (T->ndg == kKbTables) return false;
      T->diag[T->ndg] = GD.diag.p;
      T->dg[p] = T->ndg++;
    }
  }
  std::vector<double2> h(std::max<size_t>(tabs.size(), 1), make_double2(0.0, 0.0));
  for (size_t i = 0; i < tabs.size(); ++i) h[i] = make_double2(tabs[i].real(), tabs[i].imag());
  T->dc.alloc(h.size());
  T->dc.upload(h);
  return true;
}

int kblock_default_grid(const ftlm_op* op, int* grid_max) {
  if (grid_max) *grid_max = kKbMaxGrid;
  return static_cast<int>(std::min<int64_t>(kKbDefaultGrid, (op->dim + 255) / 256));
}

int kblock_bilinear_blocks(const ftlm_op* op, int R) {
  const int64_t n = op->dim * R;
  return static_cast<int>(std::min<int64_t>(kKbMaxGrid, (n + 255) / 256));
}

void kblock_bilinear(const ftlm_op* op, const KbTermSplit& T, const double2* a, const double2* b, int R,
                     double2* part, hipStream_t st) {
  const KBlock* B = static_cast<const KBlock*>(op);
  const OrbitGraph& G = *B->G;
  KbDiags dg{};
  for (int k = 0; k < T.ndg; ++k) dg.d[k] = T.diag[k];
  hipLaunchKernelGGL(k_bil_kblock, dim3(kblock_bilinear_blocks(op, R)), dim3(256), 0, st, B->rows.p, B->dim,
                     G.row_off.p, G.ent.p, B->colk.p, B->ph, T.dc.p, T.ntab, dg, T.ndg, a, b, R, part);
  check_launch("k_bil_kblock");
}

int hv_apply_kblock(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st) {
  const KBlock* B = static_cast<const KBlock*>(op);
  if (B->dim == 0) return 0;
  const HvShape sh = B->shape(R, dtype);
  const int gmax = sh.grid > 0 ? std::min(sh.grid, kKbMaxGrid) : kKbDefaultGrid;
  const int grid = static_cast<int>(std::min<int64_t>(gmax, (B->dim + 255) / 256));
  const int maxrb = sh.rb > 0 ? std::min(sh.rb, dtype == FTLM_C64 ? 16 : 8) : (dtype == FTLM_C64 ? 16 : 8);
  std::vector<std::pair<int, int>> chunks;
  for (int r0 = 0; r0 < R;) {
    int rb = 1;
    while (rb * 2 <= maxrb && r0 + rb * 2 <= R) rb *= 2;
    chunks.emplace_back(r0, rb);
    r0 += rb;
  }
  // split: the trivial lines' down hops by the full sector's row pass, then the remaining terms
  const OrbitSplit* sp = B->G->split.get();
  if (sp) hv_rows_lines(sp->sec.get(), x, y, R, dtype, xs, B->lbase.p, sp->lumask.p, sp->nlines, st);
  for (const auto& c : chunks) {
    if (dtype == FTLM_C64)
      launch_kblock_chunk<float>(B, x, y, R, c.first, c.second, xs, dotp, st, grid);
    else
      launch_kblock_chunk<double>(B, x, y, R, c.first, c.second, xs, dotp, st, grid);
  }
  check_launch("k_kblock_apply");
  return grid;
}

// ------------------------------------------------------------------ momentum stacks (host)
struct KStack : ftlm_op {
  std::shared_ptr<OrbitGraph> G;
  int nk = 0;
  std::vector<const KBlock*> blocks;  // borrowed: the caller keeps them alive
  KStackMaps maps{};
  DevBuf<uint16_t> mem;   // [nrep]
  DevBuf<uint16_t> memc;  // [nnz] mem of every term's column orbit
  DevBuf<double2> chi;    // [nk][16] Bloch phases of each block
  DevBuf<uint16_t> memc2;  // [nnz2] mem of the graph split's remaining terms' column orbits
};

const void* kblock_graph(const ftlm_op* op) {
  return op != nullptr && op->kind == kKBlock ? static_cast<const KBlock*>(op)->G.get() : nullptr;
}

std::unique_ptr<ftlm_op> kstack_build(const std::vector<const ftlm_op*>& blocks) {
  FTLM_REQUIRE(!blocks.empty() && static_cast<int>(blocks.size()) <= kMaxStackBlocks, FTLM_EINVAL,
               "a momentum stack holds 1 to 16 blocks");
  auto S = std::make_unique<KStack>();
  S->kind = kKStack;
  for (const ftlm_op* op : blocks) {
    FTLM_REQUIRE(op != nullptr && op->kind == kKBlock, FTLM_EINVAL, "a momentum stack takes momentum blocks only");
    const KBlock* B = static_cast<const KBlock*>(op);
    FTLM_REQUIRE(B->G == static_cast<const KBlock*>(blocks[0])->G, FTLM_EINVAL,
                 "the blocks of a momentum stack must share one orbit graph (create them with share=)");
    FTLM_REQUIRE(B->device == blocks[0]->device, FTLM_EINVAL, "the blocks of a momentum stack live on one device");
    FTLM_REQUIRE(B->dim >= 1, FTLM_EINVAL, "every block of a momentum stack must have a row");
    S->blocks.push_back(B);
  }
  const OrbitGraph& G = *S->blocks[0]->G;
  S->G = S->blocks[0]->G;
  S->device = blocks[0]->device;
  S->nk = static_cast<int>(blocks.size());
  S->dim = G.nrep;
  std::vector<double2> chi(static_cast<size_t>(S->nk) * kMaxOrbitSites);
  for (int j = 0; j < S->nk; ++j) {
    S->maps.map[j] = S->blocks[j]->map.p;
    for (int t = 0; t < kMaxOrbitSites; ++t) chi[static_cast<size_t>(j) * kMaxOrbitSites + t] = S->blocks[j]->ph.chi[t];
  }
  S->chi.alloc(chi.size());
  S->chi.upload(chi);
  S->mem.alloc(static_cast<size_t>(std::max<int64_t>(G.nrep, 1)));
  if (G.nrep > 0) {
    hipLaunchKernelGGL(k_kstack_mem, dim3(grid_for(G.nrep, 256)), dim3(256), 0, nullptr, G.nrep, S->nk, S->maps,
                       S->mem.p);
    check_launch("k_kstack_mem");
  }
  S->memc.alloc(static_cast<size_t>(std::max<int64_t>(G.nnz, 1)));
  if (G.nnz > 0) {
    hipLaunchKernelGGL(k_kstack_memc, dim3(grid_for(G.nnz, 256)), dim3(256), 0, nullptr, G.ent.p, G.nnz, S->mem.p,
                       S->memc.p);
    check_launch("k_kstack_memc");
  }
  if (G.split) {
    S->memc2.alloc(static_cast<size_t>(std::max<int64_t>(G.split->nnz2, 1)));
    if (G.split->nnz2 > 0) {
      hipLaunchKernelGGL(k_kstack_memc, dim3(grid_for(G.split->nnz2, 256)), dim3(256), 0, nullptr, G.split->ent2.p,
                         G.split->nnz2, S->mem.p, S->memc2.p);
      check_launch("k_kstack_memc");
    }
  }
  FTLM_HIP_CHECK(hipDeviceSynchronize());
  return S;
}

int kstack_blocks(const ftlm_op* stack) { return static_cast<const KStack*>(stack)->nk; }
KStackMaps kstack_maps(const ftlm_op* stack) { return static_cast<const KStack*>(stack)->maps; }

// Default grid of a stack apply: 2048 workgroups, 256 per XCD.
constexpr int kKsDefaultGrid = 2048;

int hv_apply_kstack(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                    hipStream_t st) {
  const KStack* S = static_cast<const KStack*>(op);
  FTLM_REQUIRE(R % S->nk == 0, FTLM_EINVAL, "a momentum-stack apply takes R = blocks * samples per block");
  if (S->dim == 0) return 0;
  const OrbitGraph& G = *S->G;
  const int rk = R / S->nk;
  const HvShape sh = S->shape(R, dtype);
  const int vpl = (dtype == FTLM_C64 && R % 2 == 0) ? 2 : 1;
  const int rpb = 256 / (R / vpl);
  const int gmax = sh.grid > 0 ? std::min(sh.grid, kKbMaxGrid) : kKsDefaultGrid;
  const int grid = static_cast<int>(std::min<int64_t>(gmax, (S->dim + rpb - 1) / rpb));
  const float2* xf = static_cast<const float2*>(x);
  float2* yf = static_cast<float2*>(y);
  // split apply: the trivial lines' down hops by the full sector's row pass, then the remaining
  // terms, added on those lines
  const OrbitSplit* sp = G.split.get();
  if (sp) hv_rows_lines(sp->sec.get(), x, y, R, dtype, xs, sp->lbase.p, sp->lumask.p, sp->nlines, st);
  const int32_t* roff = sp ? sp->row_off2.p : G.row_off.p;
  const uint2* ents = sp ? sp->ent2.p : G.ent.p;
  const uint16_t* mcs = sp ? S->memc2.p : S->memc.p;
  const uint8_t* accp = sp ? sp->triv.p : nullptr;
#define FTLM_KSTACK_ARGS S->dim, S->nk, rk, S->mem.p, mcs, roff, ents, G.diag.p, G.coef.p, G.nbond, S->chi.p, accp
  const int lanes = R / vpl;
  const bool coop = lanes >= 8 && (lanes & (lanes - 1)) == 0;
#define FTLM_KSTACK_GO(TT, VV, OJ, XP, YP)                                                                          \
  do {                                                                                                           \
    if (coop)                                                                                                    \
      hipLaunchKernelGGL((k_kstack_apply<TT, VV, OJ, true>), dim3(grid), dim3(256), 0, st, FTLM_KSTACK_ARGS, XP, YP, \
                         xs, dotp);                                                                              \
    else                                                                                                         \
      hipLaunchKernelGGL((k_kstack_apply<TT, VV, OJ, false>), dim3(grid), dim3(256), 0, st, FTLM_KSTACK_ARGS, XP, \
                         YP, xs, dotp);                                                                          \
  } while (0)
  if (dtype == FTLM_C64) {
    if (vpl == 2 && rk % 2 == 0)
      FTLM_KSTACK_GO(float, 2, true, xf, yf);
    else if (vpl == 2)
      FTLM_KSTACK_GO(float, 2, false, xf, yf);
    else
      FTLM_KSTACK_GO(float, 1, true, xf, yf);
  } else {
    FTLM_KSTACK_GO(double, 1, true, static_cast<const double2*>(x), static_cast<double2*>(y));
  }
#undef FTLM_KSTACK_GO
#undef FTLM_KSTACK_ARGS
  check_launch("k_kstack_apply");
  return grid;
}

int kstack_default_grid(const ftlm_op* op, int* grid_max) {
  if (grid_max) *grid_max = kKbMaxGrid;
  return static_cast<int>(std::min<int64_t>(kKsDefaultGrid, std::max<int64_t>(op->dim, 1)));
}

// ------------------------------------------------------------------ momentum Gram blocks
// HubbardMomentumBlock (src/hubbard_momentum_action.cpp:37-89): the momentum-K block in the
// orthonormal basis of MomentumPhiGramBasis (src/orbit_bloch_phi.cpp:226-355), applied as
// y = Phi^dag H Phi x by lifting into the full sector, applying the sector operator and projecting
// back.  The reference seeds Phi with every member of every orbit whose stabiliser has a trivial
// *bosonic* character at K (momentum_phi_seeds, k_basis.hpp:66-84, via compatible_momentum_mask,
// momentum_sector.hpp:79-97), builds the Bloch sums with the fermionic signs
// (fill_phi_orbit_bloch_from_seed, orbit_bloch_phi.hpp:36-52) and keeps the range of the Gram
// matrix.  An orbit's Bloch sum vanishes unless the fermionic signs are trivial on its stabiliser
// too, and all members of an orbit give the same column up to a phase, so the range is spanned by
// one normalised Bloch sum per orbit with eta|_Stab = 1 and chi_K|_Stab = 1: those are the columns
// used here (in the momentum block's pack order).  The reference picks another orthonormal basis
// of the same space (zheev eigenvectors, which depend on the LAPACK build); spectra, exact lnZ and
// the FTLM distribution are basis independent.  Orbits with an odd fermionic stabiliser sign drop
// out of every K, as in the reference.
struct GramBlock : ftlm_op {
  const Sector* S = nullptr;  // the full sector (not owned)
  int nt = 0;
  Phases ph{};
  DevBuf<int32_t> idx;  // [dim][nt] sector index of T_t rep; bit 31 set where eta(rep, t) = -1
  DevBuf<double> cs;    // [dim] 1 / sqrt(|G| |Stab|)
  mutable DevBuf<unsigned char> xf, yf;  // full-sector scratch blocks
  mutable size_t xf_layout = 0;           // bytes per state of xf's last clear (0: not cleared)
};

namespace {

__global__ void k_gram_flag(const int32_t* rows, int64_t n, const uint16_t* stab, const uint16_t* eta,
                            uint8_t* flag) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t g = rows[r];
  flag[r] = (stab[g] & eta[g]) == 0 ? 1 : 0;
}

__global__ void k_gram_images(const int32_t* rows, const int32_t* sel, int64_t dk, const uint32_t* pack,
                              const uint16_t* stab, LatDev L, uint32_t* img, uint8_t* neg, double* cs) {
  const int64_t a = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= dk) return;
  const int32_t g = rows[sel[a]];
  const uint32_t rep = pack[g], u = rep >> 16, d = rep & 0xFFFFu;
  cs[a] = 1.0 / sqrt(static_cast<double>(L.nt) * static_cast<double>(__popc(static_cast<uint32_t>(stab[g]))));
  for (int t = 0; t < L.nt; ++t) {
    img[a * L.nt + t] = pack_state(tr_bits(u, L, t), tr_bits(d, L, t));
    neg[a * L.nt + t] = tr_parity(u, d, L, t) ? 1 : 0;
  }
}

__global__ void k_gram_sign(int32_t* idx, const uint8_t* neg, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && neg[i]) idx[i] = static_cast<int32_t>(static_cast<uint32_t>(idx[i]) | 0x80000000u);
}

// xf[T_t rep] = |Stab| c eta_t conj(chi_t) x_a for the vectors r0 .. r0+RB of the block of RT:
// the Bloch sum reaches each member once per stabiliser element, always with the same coefficient,
// so a member's amplitude is |Stab| c eta conj(chi) = sqrt(|Stab| / |G|) eta conj(chi) and every
// write of it stores the whole amplitude.  xf holds SB >= RB vectors per state.
template <typename T>
__global__ __launch_bounds__(256) void k_gram_lift(int64_t n, int RT, int r0, int RB, int SB, int nt,
                                                   const int32_t* __restrict__ idx, const double* __restrict__ cs,
                                                   Phases ph, const typename Cx<T>::v2* __restrict__ x,
                                                   typename Cx<T>::v2* __restrict__ xf) {
  using V2 = typename Cx<T>::v2;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t a = f / RB;
    const int rr = static_cast<int>(f - a * RB);
    const V2 v = x[a * RT + r0 + rr];
    const double vr = static_cast<double>(v.x), vi = static_cast<double>(v.y);
    const double c = 1.0 / (static_cast<double>(nt) * cs[a]);  // |Stab| c, with c = 1 / sqrt(|G| |Stab|)
    for (int t = 0; t < nt; ++t) {
      const int32_t e = idx[a * nt + t];
      const double sc = e < 0 ? -c : c;
      const double kr = sc * ph.chi[t].x, ki = -sc * ph.chi[t].y;  // |Stab| c eta conj(chi)
      V2 o;
      o.x = static_cast<T>(kr * vr - ki * vi);
      o.y = static_cast<T>(kr * vi + ki * vr);
      xf[static_cast<int64_t>(e & 0x7FFFFFFF) * SB + rr] = o;
    }
  }
}

// y_a = xs_r sum_t c eta_t chi_t yf[T_t rep] (= Phi^dag yf) for r = r0 .. r0+RB, with <xs x | y>
// partials per block in the layout of the sector applies.  The block size is a multiple of RB;
// yf holds SB >= RB vectors per state.
template <typename T>
__global__ __launch_bounds__(256) void k_gram_project(int64_t n, int RT, int r0, int RB, int SB, int nt,
                                                      const int32_t* __restrict__ idx, const double* __restrict__ cs,
                                                      Phases ph, const typename Cx<T>::v2* __restrict__ yf,
                                                      typename Cx<T>::v2* __restrict__ y,
                                                      const typename Cx<T>::v2* __restrict__ x,
                                                      const double* __restrict__ xs, double2* __restrict__ dotp) {
  using V2 = typename Cx<T>::v2;
  const int nthr = blockDim.x, rr = threadIdx.x % RB;
  const double sx = xs ? xs[r0 + rr] : 1.0;
  double dre = 0.0, dim = 0.0;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * nthr + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * nthr) {
    const int64_t a = f / RB;
    const double c = cs[a];
    double ar = 0.0, ai = 0.0;
    for (int t = 0; t < nt; ++t) {
      const int32_t e = idx[a * nt + t];
      const double sc = e < 0 ? -c : c;
      const double kr = sc * ph.chi[t].x, ki = sc * ph.chi[t].y;  // c eta chi
      const V2 w = yf[static_cast<int64_t>(e & 0x7FFFFFFF) * SB + rr];
      const double wr = static_cast<double>(w.x), wi = static_cast<double>(w.y);
      ar += kr * wr - ki * wi;
      ai += kr * wi + ki * wr;
    }
    V2 o;
    o.x = static_cast<T>(sx * ar);
    o.y = static_cast<T>(sx * ai);
    const int64_t g = a * RT + r0 + rr;
    y[g] = o;
    if (dotp) {
      const V2 q = x[g];
      const double qr = sx * static_cast<double>(q.x), qi = sx * static_cast<double>(q.y);
      dre += qr * static_cast<double>(o.x) + qi * static_cast<double>(o.y);
      dim += qr * static_cast<double>(o.y) - qi * static_cast<double>(o.x);
    }
  }
  if (dotp) {
    __shared__ double red[2][256];
    red[0][threadIdx.x] = dre;
    red[1][threadIdx.x] = dim;
    __syncthreads();
    if (threadIdx.x < RB) {
      double u = 0.0, v = 0.0;
      for (int t = threadIdx.x; t < nthr; t += RB) {
        u += red[0][t];
        v += red[1][t];
      }
      dotp[static_cast<int64_t>(blockIdx.x) * RT + r0 + threadIdx.x] = make_double2(u, v);
    }
  }
}

__global__ void k_gram_zero_d2(double2* dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = make_double2(0.0, 0.0);
}

constexpr int kGramGrid = 2048;
constexpr int kGramChunk = 32;  // vectors per lift / sector apply / projection round

}  // namespace

// Vectors go through in chunks of at most kGramChunk.  The two full-sector scratch blocks hold SB
// = min(R, 32) vectors per state for every chunk of an apply (a shorter last chunk leaves its
// spare vectors unused).  The lift writes the block's member states only, so the others must read
// zero: the scratch is cleared whenever its per-state layout changes.
int hv_apply_gram(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
                  hipStream_t st) {
  const GramBlock* B = static_cast<const GramBlock*>(op);
  if (B->dim == 0) return 0;
  const size_t es = dtype == FTLM_C64 ? sizeof(float2) : sizeof(double2);
  const int SB = std::min(R, kGramChunk);
  const size_t fb = static_cast<size_t>(B->S->dim) * SB * es;
  if (B->xf.n < fb) {
    B->xf.alloc(fb);
    B->yf.alloc(fb);
    B->xf_layout = 0;
  }
  if (B->xf_layout != SB * es) {
    FTLM_HIP_CHECK(hipMemsetAsync(B->xf.p, 0, fb, st));
    B->xf_layout = SB * es;
  }
  if (dotp && R > SB) {  // partial blocks a chunk does not write must read as zero
    const int64_t nz = static_cast<int64_t>(kGramGrid) * R;
    hipLaunchKernelGGL(k_gram_zero_d2, dim3(static_cast<unsigned>(std::min<int64_t>(1024, (nz + 255) / 256))),
                       dim3(256), 0, st, dotp, nz);
  }
  int nblk = 0;
  for (int r0 = 0; r0 < R; r0 += SB) {
    const int RB = std::min(SB, R - r0);
    const int64_t n = B->dim * RB;
    const int thr = (256 / RB) * RB;
    const int grid = static_cast<int>(std::min<int64_t>(kGramGrid, (n + thr - 1) / thr));
    nblk = std::max(nblk, grid);
    if (dtype == FTLM_C64)
      hipLaunchKernelGGL(k_gram_lift<float>, dim3(grid), dim3(256), 0, st, n, R, r0, RB, SB, B->nt, B->idx.p,
                         B->cs.p, B->ph, static_cast<const float2*>(x),
                         static_cast<float2*>(static_cast<void*>(B->xf.p)));
    else
      hipLaunchKernelGGL(k_gram_lift<double>, dim3(grid), dim3(256), 0, st, n, R, r0, RB, SB, B->nt, B->idx.p,
                         B->cs.p, B->ph, static_cast<const double2*>(x),
                         static_cast<double2*>(static_cast<void*>(B->xf.p)));
    check_launch("k_gram_lift");
    hv_apply(B->S, B->xf.p, B->yf.p, SB, dtype, nullptr, nullptr, st);
    if (dtype == FTLM_C64)
      hipLaunchKernelGGL(k_gram_project<float>, dim3(grid), dim3(thr), 0, st, n, R, r0, RB, SB, B->nt, B->idx.p,
                         B->cs.p, B->ph, static_cast<const float2*>(static_cast<const void*>(B->yf.p)),
                         static_cast<float2*>(y), static_cast<const float2*>(x), xs, dotp);
    else
      hipLaunchKernelGGL(k_gram_project<double>, dim3(grid), dim3(thr), 0, st, n, R, r0, RB, SB, B->nt, B->idx.p,
                         B->cs.p, B->ph, static_cast<const double2*>(static_cast<const void*>(B->yf.p)),
                         static_cast<double2*>(y), static_cast<const double2*>(x), xs, dotp);
    check_launch("k_gram_project");
  }
  return nblk;
}

}  // namespace ftlm_hip

using namespace ftlm_hip;

namespace {

void check_lattice(int Lx, int Ly) {
  FTLM_REQUIRE(Lx >= 1 && Ly >= 1, FTLM_EINVAL, "lattice extents must be positive");
  FTLM_REQUIRE(Lx * Ly <= kMaxOrbitSites, FTLM_EINVAL, "OrbitKBlockBasis::build: n_sites must be <= 16 (RawState bit width)");
}

const KBlock* as_kblock(const ftlm_op_t* op) {
  FTLM_REQUIRE(op != nullptr && op->kind == kKBlock, FTLM_EINVAL, "not a k-block operator");
  return static_cast<const KBlock*>(op);
}

}  // namespace

namespace {

void kstack_copy(const ftlm_op_t* stack, int j, const void* src, int rk, int dtype, void* dst, void* stream,
                 bool scatter) {
  FTLM_REQUIRE(stack != nullptr && stack->kind == kKStack, FTLM_EINVAL, "not a momentum-stack operator");
  FTLM_REQUIRE(src != nullptr && dst != nullptr, FTLM_EINVAL, "null argument");
  FTLM_REQUIRE(dtype == FTLM_C128 || dtype == FTLM_C64, FTLM_EINVAL, "dtype must be FTLM_C128 or FTLM_C64");
  const KStack* S = static_cast<const KStack*>(stack);
  FTLM_REQUIRE(j >= 0 && j < S->nk, FTLM_ERANGE, "block index outside the stack");
  FTLM_REQUIRE(rk >= 1 && S->nk * rk <= 32, FTLM_EINVAL, "blocks * samples per block must be in [1, 32]");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = S->dim * rk;
  if (n == 0) return;
  const unsigned g = static_cast<unsigned>(std::min<int64_t>(8192, (n + 255) / 256));
  const int W = S->nk * rk;
  const int32_t* map = S->maps.map[j];
  if (dtype == FTLM_C64) {
    if (scatter)
      hipLaunchKernelGGL(k_kstack_scatter<float>, dim3(g), dim3(256), 0, st, S->dim, W, j, rk, map,
                         static_cast<const float2*>(src), static_cast<float2*>(dst));
    else
      hipLaunchKernelGGL(k_kstack_gather<float>, dim3(g), dim3(256), 0, st, S->dim, W, j, rk, map,
                         static_cast<const float2*>(src), static_cast<float2*>(dst));
  } else {
    if (scatter)
      hipLaunchKernelGGL(k_kstack_scatter<double>, dim3(g), dim3(256), 0, st, S->dim, W, j, rk, map,
                         static_cast<const double2*>(src), static_cast<double2*>(dst));
    else
      hipLaunchKernelGGL(k_kstack_gather<double>, dim3(g), dim3(256), 0, st, S->dim, W, j, rk, map,
                         static_cast<const double2*>(src), static_cast<double2*>(dst));
  }
  check_launch(scatter ? "k_kstack_scatter" : "k_kstack_gather");
}
}  // namespace

extern "C" {

int ftlm_hip_kblock_create(int Lx, int Ly, int n_up, int n_dn, int kx, int ky, double tol, const ftlm_hop_t* hops,
                           int n_hops, const ftlm_pair_t* pairs, int n_pairs, double U, double V,
                           const ftlm_op_t* share, ftlm_op_t** out) {
  return guarded([&] {
    FTLM_REQUIRE(out != nullptr, FTLM_EINVAL, "null output handle");
    *out = nullptr;
    check_lattice(Lx, Ly);
    const int n = Lx * Ly;
    FTLM_REQUIRE(n_up >= 0 && n_up <= n && n_dn >= 0 && n_dn <= n, FTLM_EINVAL,
                 "FockBasis: particle numbers must be in [0, n_sites]");
    FTLM_REQUIRE(kx >= 0 && kx < Lx && ky >= 0 && ky < Ly, FTLM_EINVAL, "momentum (kx, ky) outside the lattice");
    FTLM_REQUIRE(n_hops >= 0 && (n_hops == 0 || hops != nullptr), FTLM_EINVAL, "null hop list");
    FTLM_REQUIRE(n_pairs >= 0 && (n_pairs == 0 || pairs != nullptr), FTLM_EINVAL, "null pair list");
    int dev = 0;
    FTLM_HIP_CHECK(hipGetDevice(&dev));
    auto B = std::make_unique<KBlock>();
    B->kind = kKBlock;
    B->device = dev;
    B->kx = kx;
    B->ky = ky;
    B->tol = tol;
    if (share != nullptr) {
      const KBlock* S = as_kblock(share);
      FTLM_REQUIRE(S->G->same_model(Lx, Ly, n_up, n_dn, U, V, hops, n_hops, pairs, n_pairs), FTLM_EINVAL,
                   "share: the operator belongs to another sector or Hamiltonian");
      FTLM_REQUIRE(S->device == dev, FTLM_EINVAL, "share: the operator lives on another device");
      B->G = S->G;
    } else {
      B->G = build_orbit_graph(Lx, Ly, n_up, n_dn, hops, n_hops, pairs, n_pairs, U, V);
    }
    build_kblock(B.get(), tol);
    *out = B.release();
  });
}

int ftlm_hip_kstack_create(ftlm_op_t* const* blocks, int n_blocks, ftlm_op_t** out) {
  return guarded([&] {
    FTLM_REQUIRE(out != nullptr && blocks != nullptr, FTLM_EINVAL, "null argument");
    *out = nullptr;
    FTLM_REQUIRE(n_blocks >= 1 && n_blocks <= kMaxStackBlocks, FTLM_EINVAL, "a momentum stack holds 1 to 16 blocks");
    std::vector<const ftlm_op*> v(blocks, blocks + n_blocks);
    FTLM_REQUIRE(v[0] != nullptr, FTLM_EINVAL, "null block");
    FTLM_HIP_CHECK(hipSetDevice(v[0]->device));
    std::unique_ptr<ftlm_op> S = kstack_build(v);
    *out = S.release();
  });
}


int ftlm_hip_kstack_scatter(const ftlm_op_t* stack, int j, const void* xb, int rk, int dtype, void* xs,
                            void* stream) {
  return guarded([&] { kstack_copy(stack, j, xb, rk, dtype, xs, stream, true); });
}

int ftlm_hip_kstack_gather(const ftlm_op_t* stack, int j, const void* xs, int rk, int dtype, void* xb,
                           void* stream) {
  return guarded([&] { kstack_copy(stack, j, xs, rk, dtype, xb, stream, false); });
}

int ftlm_hip_kgram_create(const ftlm_op_t* sector, const ftlm_op_t* kblock, ftlm_op_t** out) {
  return guarded([&] {
    FTLM_REQUIRE(out != nullptr, FTLM_EINVAL, "null output handle");
    *out = nullptr;
    FTLM_REQUIRE(sector != nullptr && sector->kind == kSector, FTLM_EINVAL, "not a full-sector operator");
    const KBlock* K = as_kblock(kblock);
    const Sector* S = static_cast<const Sector*>(sector);
    const OrbitGraph& G = *K->G;
    FTLM_REQUIRE(S->n_sites == G.lx * G.ly && S->n_up == G.n_up && S->n_dn == G.n_dn, FTLM_EINVAL,
                 "the sector and the momentum block belong to different (lattice, N_up, N_dn)");
    FTLM_REQUIRE(S->dim < (int64_t{1} << 31), FTLM_EUNSUPPORTED, "sector too large for 32-bit state indices");
    FTLM_HIP_CHECK(hipSetDevice(K->device));
    auto B = std::make_unique<GramBlock>();
    B->kind = kGram;
    B->device = K->device;
    B->S = S;
    B->nt = G.L.nt;
    B->ph = K->ph;
    const hipStream_t st = nullptr;
    const int64_t nk = K->dim;
    int32_t dk = 0;
    DevBuf<int32_t> sel(std::max<int64_t>(nk, 1)), nsel(1);
    if (nk > 0) {
      DevBuf<uint8_t> flag(static_cast<size_t>(nk));
      hipLaunchKernelGGL(k_gram_flag, dim3(grid_for(nk, 256)), dim3(256), 0, st, K->rows.p, nk, G.stab.p, G.eta.p,
                         flag.p);
      check_launch("k_gram_flag");
      hipcub::CountingInputIterator<int32_t> it(0);
      const int nn = static_cast<int>(nk);
      cub_call([&](void* t, size_t& b, hipStream_t s) {
        return hipcub::DeviceSelect::Flagged(t, b, it, flag.p, sel.p, nsel.p, nn, s);
      }, st);
      FTLM_HIP_CHECK(hipMemcpy(&dk, nsel.p, sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    B->dim = dk;
    const int64_t nimg = static_cast<int64_t>(dk) * B->nt;
    B->idx.alloc(static_cast<size_t>(std::max<int64_t>(nimg, 1)));
    B->cs.alloc(static_cast<size_t>(std::max<int32_t>(dk, 1)));
    if (dk > 0) {
      DevBuf<uint32_t> img(static_cast<size_t>(nimg));
      DevBuf<uint8_t> neg(static_cast<size_t>(nimg));
      hipLaunchKernelGGL(k_gram_images, dim3(grid_for(dk, 256)), dim3(256), 0, st, K->rows.p, sel.p,
                         static_cast<int64_t>(dk), G.pack.p, G.stab.p, G.L, img.p, neg.p, B->cs.p);
      check_launch("k_gram_images");
      sector_rank_packs(S, img.p, nimg, B->idx.p, st);
      hipLaunchKernelGGL(k_gram_sign, dim3(grid_for(nimg, 256)), dim3(256), 0, st, B->idx.p, neg.p, nimg);
      check_launch("k_gram_sign");
    }
    FTLM_HIP_CHECK(hipDeviceSynchronize());
    *out = B.release();
  });
}

int ftlm_hip_kblock_info(const ftlm_op_t* op, int* kx, int* ky, int64_t* n_orbits, int64_t* nnz) {
  return guarded([&] {
    const KBlock* B = as_kblock(op);
    if (kx) *kx = B->kx;
    if (ky) *ky = B->ky;
    if (n_orbits) *n_orbits = B->G->nrep;
    if (nnz) *nnz = B->G->nnz;
  });
}

int ftlm_hip_kblock_entries(const ftlm_op_t* op, int64_t first, int64_t count, uint32_t* rep, int* orbit_size,
                            int* stab_size, double* A) {
  return guarded([&] {
    const KBlock* B = as_kblock(op);
    FTLM_REQUIRE(first >= 0 && count >= 0 && first + count <= B->dim, FTLM_ERANGE, "entry range outside the block");
    if (count == 0) return;
    FTLM_HIP_CHECK(hipSetDevice(B->device));
    const OrbitGraph& G = *B->G;
    const size_t n = static_cast<size_t>(count);
    DevBuf<uint32_t> dr(n);
    DevBuf<int> dos(n), dss(n);
    DevBuf<double2> da(n);
    hipLaunchKernelGGL(k_kblock_entries, dim3(grid_for(count, 256)), dim3(256), 0, 0, B->rows.p, first, count,
                       G.pack.p, G.stab.p, B->A.p, G.L.nt, dr.p, dos.p, dss.p, da.p);
    check_launch("k_kblock_entries");
    if (rep) FTLM_HIP_CHECK(hipMemcpy(rep, dr.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (orbit_size) FTLM_HIP_CHECK(hipMemcpy(orbit_size, dos.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (stab_size) FTLM_HIP_CHECK(hipMemcpy(stab_size, dss.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (A) FTLM_HIP_CHECK(hipMemcpy(A, da.p, sizeof(double2) * n, hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_kblock_orbit_index(const ftlm_op_t* op, int64_t first, int64_t count, int64_t* out) {
  return guarded([&] {
    const KBlock* B = as_kblock(op);
    FTLM_REQUIRE(first >= 0 && count >= 0 && first + count <= B->dim, FTLM_ERANGE, "entry range outside the block");
    FTLM_REQUIRE(count == 0 || out != nullptr, FTLM_EINVAL, "null output");
    if (count == 0) return;
    std::vector<int32_t> rows(static_cast<size_t>(count));
    FTLM_HIP_CHECK(hipMemcpy(rows.data(), B->rows.p + first, sizeof(int32_t) * count, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < count; ++i) out[i] = rows[static_cast<size_t>(i)];
  });
}

int ftlm_hip_kblock_index_of_rep(const ftlm_op_t* op, const uint32_t* reps, int64_t count, int64_t* out) {
  return guarded([&] {
    const KBlock* B = as_kblock(op);
    FTLM_REQUIRE(count >= 0 && (count == 0 || (reps && out)), FTLM_EINVAL, "null argument");
    if (count == 0) return;
    FTLM_HIP_CHECK(hipSetDevice(B->device));
    DevBuf<uint32_t> dr(static_cast<size_t>(count));
    DevBuf<int64_t> dout(static_cast<size_t>(count));
    dr.upload(reps, static_cast<size_t>(count));
    hipLaunchKernelGGL(k_kblock_index_of_rep, dim3(grid_for(count, 256)), dim3(256), 0, 0, dr.p, count, B->G->pack.p,
                       B->G->nrep, B->map.p, dout.p);
    check_launch("k_kblock_index_of_rep");
    FTLM_HIP_CHECK(hipMemcpy(out, dout.p, sizeof(int64_t) * count, hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_kblock_map(const ftlm_op_t* op, const uint32_t* states, int64_t count, int64_t* index, int* dx, int* dy,
                        double* eta) {
  return guarded([&] {
    const KBlock* B = as_kblock(op);
    FTLM_REQUIRE(count >= 0 && (count == 0 || (states && index && dx && dy && eta)), FTLM_EINVAL, "null argument");
    if (count == 0) return;
    FTLM_HIP_CHECK(hipSetDevice(B->device));
    const size_t n = static_cast<size_t>(count);
    DevBuf<uint32_t> ds(n);
    DevBuf<int64_t> di(n);
    DevBuf<int> ddx(n), ddy(n);
    DevBuf<double> de(n);
    ds.upload(states, n);
    hipLaunchKernelGGL(k_kblock_map_states, dim3(grid_for(count, 256)), dim3(256), 0, 0, ds.p, count, B->G->pack.p,
                       B->G->nrep, B->map.p, B->G->L, di.p, ddx.p, ddy.p, de.p);
    check_launch("k_kblock_map_states");
    FTLM_HIP_CHECK(hipMemcpy(index, di.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    FTLM_HIP_CHECK(hipMemcpy(dx, ddx.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    FTLM_HIP_CHECK(hipMemcpy(dy, ddy.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    FTLM_HIP_CHECK(hipMemcpy(eta, de.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_translate_states(int Lx, int Ly, const uint32_t* states, int64_t count, int dx, int dy, uint32_t* out,
                              double* sign) {
  return guarded([&] {
    check_lattice(Lx, Ly);
    FTLM_REQUIRE(count >= 0 && (count == 0 || states != nullptr), FTLM_EINVAL, "null argument");
    if (count == 0) return;
    const int tx = ((dx % Lx) + Lx) % Lx, ty = ((dy % Ly) + Ly) % Ly;
    const LatDev L = make_lat(Lx, Ly);
    const size_t n = static_cast<size_t>(count);
    DevBuf<uint32_t> ds(n), dout(n);
    DevBuf<double> dsg(n);
    ds.upload(states, n);
    hipLaunchKernelGGL(k_translate_states, dim3(grid_for(count, 256)), dim3(256), 0, 0, ds.p, count, L,
                       ty * Lx + tx, dout.p, dsg.p);
    check_launch("k_translate_states");
    if (out) FTLM_HIP_CHECK(hipMemcpy(out, dout.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (sign) FTLM_HIP_CHECK(hipMemcpy(sign, dsg.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_canonical_states(int Lx, int Ly, const uint32_t* states, int64_t count, uint32_t* canon,
                              int* stab_size) {
  return guarded([&] {
    check_lattice(Lx, Ly);
    FTLM_REQUIRE(count >= 0 && (count == 0 || states != nullptr), FTLM_EINVAL, "null argument");
    if (count == 0) return;
    const LatDev L = make_lat(Lx, Ly);
    const size_t n = static_cast<size_t>(count);
    DevBuf<uint32_t> ds(n), dc(n);
    DevBuf<int> dss(n);
    ds.upload(states, n);
    hipLaunchKernelGGL(k_canonical_states, dim3(grid_for(count, 256)), dim3(256), 0, 0, ds.p, count, L, dc.p, dss.p);
    check_launch("k_canonical_states");
    if (canon) FTLM_HIP_CHECK(hipMemcpy(canon, dc.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    if (stab_size) FTLM_HIP_CHECK(hipMemcpy(stab_size, dss.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_bloch_phase(int Lx, int Ly, int kx, int ky, int dx, int dy, double* re, double* im) {
  return guarded([&] {
    FTLM_REQUIRE(Lx >= 1 && Ly >= 1 && re && im, FTLM_EINVAL, "bad argument");
    const double pi = std::acos(-1.0);
    const double th = -2.0 * pi *
                      (static_cast<double>(kx * dx) / static_cast<double>(Lx) +
                       static_cast<double>(ky * dy) / static_cast<double>(Ly));
    *re = std::cos(th);
    *im = std::sin(th);
  });
}

}  // extern "C"
