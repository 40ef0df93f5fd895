// ftlm_oracle.cpp -- CPU restatement of the reference FTLM hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (ftlm_extended_hubbard_amd/) may link, load
// or call this file.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker / the timed CPU baseline.
//
// Each function restates one reference routine (paths relative to the reference checkout) and is
// written so that the floating-point operation order matches the reference where parity is
// bit-level (geometry, push-form H*v, Lanczos, capped Jacobi).  Pinning of every function is
// listed in the table at the bottom of this header comment; none is "parity unpinned".
//
//   orc_geometry               src/lattice.cpp:9-110
//   orc_basis_*                src/fock_basis.cpp:7-51
//   orc_apply                  src/hubbard_hamiltonian.cpp:24-108
//   orc_lanczos                src/lanczos.cpp:125-209 (+ optional supplied start vector)
//   orc_tridiagonal_extrema    src/lanczos.cpp:38-123 (120-rotation capped Jacobi)
//   orc_quadrature             src/ftlm_thermo.cpp:17-146 (max(3000,120n) capped Jacobi)
//   orc_ftlm_log_partition_*   src/ftlm_thermo.cpp:150-256
//   orc_gc_density             benchmarks/bench_ftlm_nmu_rect.cpp:186-219
//   orc_orbit_*                src/orbit_k_block_basis.cpp:56-220, include/ftlm/symmetry/*.hpp
//   orc_orbit_apply_rows       src/orbit_k_block_basis.cpp:127-220 restricted to chosen rows
//   orc_orbit_lift             include/ftlm/symmetry/orbit_bloch_phi.hpp:35-51 (Phi_K v)
//   orc_exact_log_partition    benchmarks/bench_ftlm_nmu_4x4_k.cpp:249-336
//   orc_dense_extrema          benchmarks/bench_2x2.cpp:28-93,157-178 (1600-rotation Jacobi)
//   orc_tridiag_eig_ql         converged symmetric tridiagonal eigensolver (implicit QL with
//                              Wilkinson shift, textbook), used to state converged values.
//
// Pinning (what checks each function; see tests/test_oracle.py):
//   geometry / basis / rank / apply / lanczos / extrema / quadrature / ftlm / orbit:
//       bit-exact against oracle/_ref (the reference sources compiled here) through
//       tests/golden fixtures; basis against brute force; apply against the published 2x2
//       spectra (benchmarks/data/eigs_full_vs_k_2x2.tsv) and numpy eigvalsh.
//   gc_density: published 2x2 ED n(mu) (benchmarks/data/n_vs_mu_beta20.tsv).
//   exact_log_partition: numpy eigvalsh of the dense block (closed form ln sum e^{-beta E}).
//   tridiag_eig_ql: numpy eigh_tridiagonal / eigvalsh.
//   orbit dims: Sigma_K dim_K = dim_full identity, reference test values 12 and 66.
//   orbit_apply_rows: equal to the full block push orc_orbit_apply on 2x2/3x2/4x2/4x3 blocks
//       (flux included) and refuses a representative list that disagrees with the basis.
//   orbit_lift: an isometry with H Phi_K v = Phi_K H_K v (orc_apply, orc_orbit_apply) at zero
//       flux, and not with the conjugated phase convention.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

using cplx = std::complex<double>;

struct Hop {
  int from;
  int to;
  int spin;
  cplx coeff;
};

struct Pair {
  int i;
  int j;
};

struct Model {
  int Lx, Ly;
  double t, tp, U, V, phx, phy;
};

Model model_from(const double* p) {
  Model m;
  m.Lx = static_cast<int>(p[0]);
  m.Ly = static_cast<int>(p[1]);
  m.t = p[2];
  m.tp = p[3];
  m.U = p[4];
  m.V = p[5];
  m.phx = p[6];
  m.phy = p[7];
  return m;
}

// ---------------------------------------------------------------- geometry (src/lattice.cpp)

void push_bond(std::vector<Hop>* hops, int a, int b, cplx phase, double ts) {
  // src/lattice.cpp:13-24: both directions, both spins, skipped for |t| < 1e-15.
  if (std::abs(ts) < 1e-15) return;
  const cplx fwd = -ts * phase;
  const cplx bwd = -ts * std::conj(phase);
  for (int s = 0; s < 2; ++s) {
    hops->push_back(Hop{a, b, s, fwd});
    hops->push_back(Hop{b, a, s, bwd});
  }
}

void geometry(const Model& m, std::vector<Hop>* hops, std::vector<Pair>* pairs) {
  hops->clear();
  pairs->clear();
  auto nn_pair = [&](int a, int b) {
    if (a == b) return;
    if (a > b) std::swap(a, b);
    pairs->push_back(Pair{a, b});
  };
  for (int y = 0; y < m.Ly; ++y) {
    for (int x = 0; x < m.Lx; ++x) {
      const int i = y * m.Lx + x;
      if (m.Lx > 1) {  // east neighbour, Peierls phase on the x wrap
        int nx = x + 1, wx = 0;
        if (nx >= m.Lx) { nx = 0; wx = 1; }
        const int j = y * m.Lx + nx;
        if (i != j) {
          nn_pair(i, j);
          push_bond(hops, i, j, std::polar(1.0, wx * m.phx + 0 * m.phy), m.t);
        }
      }
      if (m.Ly > 1) {  // north neighbour
        int ny = y + 1, wy = 0;
        if (ny >= m.Ly) { ny = 0; wy = 1; }
        const int j = ny * m.Lx + x;
        if (i != j) {
          nn_pair(i, j);
          push_bond(hops, i, j, std::polar(1.0, 0 * m.phx + wy * m.phy), m.t);
        }
      }
      const int dirs[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
      for (const auto& d : dirs) {  // all four diagonals from every site (src/lattice.cpp:80-107)
        int nx = x + d[0], ny = y + d[1], wx = 0, wy = 0;
        if (nx >= m.Lx) { nx -= m.Lx; wx = 1; } else if (nx < 0) { nx += m.Lx; wx = -1; }
        if (ny >= m.Ly) { ny -= m.Ly; wy = 1; } else if (ny < 0) { ny += m.Ly; wy = -1; }
        const int j = ny * m.Lx + nx;
        if (i != j) push_bond(hops, i, j, std::polar(1.0, wx * m.phx + wy * m.phy), m.tp);
      }
    }
  }
}

// ---------------------------------------------------------------- basis (src/fock_basis.cpp)

void combos(int n, int k, int first, uint64_t acc, std::vector<uint64_t>* out) {
  if (k == 0) {
    out->push_back(acc);
    return;
  }
  for (int p = first; p <= n - k; ++p) combos(n, k - 1, p + 1, acc | (uint64_t{1} << p), out);
}

struct Basis {
  int n = 0, nu = 0, nd = 0;
  std::vector<uint64_t> ups, dns;  // per-spin strings in lexicographic order
  std::unordered_map<uint64_t, long> index;  // (u << n) | d -> iu * Ddn + id
  bool ok = false;
  long dim() const { return static_cast<long>(ups.size() * dns.size()); }
  uint64_t up(long i) const { return ups[static_cast<size_t>(i / static_cast<long>(dns.size()))]; }
  uint64_t dn(long i) const { return dns[static_cast<size_t>(i % static_cast<long>(dns.size()))]; }
  long index_of(uint64_t u, uint64_t d) const {
    auto it = index.find((u << n) | d);
    return it == index.end() ? -1 : it->second;
  }
};

bool make_basis(int n, int nu, int nd, Basis* b, bool with_index = true) {
  b->ok = false;
  if (n <= 0 || n > 32) return false;
  if (nu < 0 || nd < 0 || nu > n || nd > n) return false;
  b->n = n;
  b->nu = nu;
  b->nd = nd;
  b->ups.clear();
  b->dns.clear();
  combos(n, nu, 0, 0, &b->ups);
  combos(n, nd, 0, 0, &b->dns);
  b->index.clear();
  if (with_index) {
    long idx = 0;
    for (uint64_t u : b->ups)
      for (uint64_t d : b->dns) b->index[(u << n) | d] = idx++;
  }
  b->ok = true;
  return true;
}

// Closed-form lexicographic rank (used only by the column-slice CPU baseline; identical results
// to the hash lookup, checked in tests).
struct Ranker {
  int n = 0, k = 0;
  std::vector<std::vector<long>> C;
  void init(int nn, int kk) {
    n = nn;
    k = kk;
    C.assign(static_cast<size_t>(n + 1), std::vector<long>(static_cast<size_t>(n + 1), 0));
    for (int a = 0; a <= n; ++a) {
      C[a][0] = 1;
      for (int b = 1; b <= a; ++b) C[a][b] = C[a - 1][b - 1] + (b <= a - 1 ? C[a - 1][b] : 0);
    }
  }
  long binom(int a, int b) const { return (a < 0 || b < 0 || b > a) ? 0 : C[a][b]; }
  long rank(uint64_t mask) const {
    long r = 0;
    int prev = -1, i = 0;
    for (int p = 0; p < n; ++p) {
      if (!((mask >> p) & 1)) continue;
      ++i;
      for (int q = prev + 1; q < p; ++q) r += binom(n - 1 - q, k - i);
      prev = p;
    }
    return r;
  }
};

// ---------------------------------------------------------------- apply (src/hubbard_hamiltonian.cpp)

int between(int ma, int mb, uint64_t up, uint64_t dn) {
  if (ma > mb) std::swap(ma, mb);
  int c = 0;
  for (int m = ma + 1; m < mb; ++m) {
    const uint64_t s = (m % 2 == 0) ? up : dn;
    if ((s >> (m / 2)) & 1) ++c;
  }
  return c;
}

int occ(uint64_t u, uint64_t d, int site) {
  return static_cast<int>(((u >> site) & 1) + ((d >> site) & 1));
}

template <class IndexFn>
void apply_cols(const Model& m, const Basis& b, const std::vector<Hop>& hops,
                const std::vector<Pair>& pairs, const cplx* x, cplx* y, long c0, long c1,
                IndexFn&& index_of) {
  for (long col = c0; col < c1; ++col) {
    const uint64_t u = b.up(col), d = b.dn(col);
    const cplx vc = x[col];
    if (std::norm(vc) < 1e-28) continue;
    cplx dg(static_cast<double>(__builtin_popcountll(u & d)) * m.U, 0.0);
    for (const Pair& pr : pairs) dg += m.V * static_cast<double>(occ(u, d, pr.i) * occ(u, d, pr.j));
    y[col] += dg * vc;
    for (const Hop& h : hops) {
      uint64_t u2 = u, d2 = d;
      const uint64_t src = h.spin == 0 ? u : d;
      if (!((src >> h.from) & 1)) continue;
      if ((src >> h.to) & 1) continue;
      const uint64_t flip = (uint64_t{1} << h.from) ^ (uint64_t{1} << h.to);
      if (h.spin == 0) u2 ^= flip; else d2 ^= flip;
      const int nb = between(2 * h.from + h.spin, 2 * h.to + h.spin, u, d);
      const double sg = (nb & 1) ? -1.0 : 1.0;
      const long row = index_of(u2, d2);
      if (row < 0) continue;
      y[row] += (sg * h.coeff) * vc;
    }
  }
}

void apply_full(const Model& m, const Basis& b, const std::vector<Hop>& hops,
                const std::vector<Pair>& pairs, const cplx* x, cplx* y) {
  const long dim = b.dim();
  std::fill(y, y + dim, cplx(0.0, 0.0));
  apply_cols(m, b, hops, pairs, x, y, 0, dim,
             [&](uint64_t u, uint64_t d) { return b.index_of(u, d); });
}

// ---------------------------------------------------------------- Lanczos (src/lanczos.cpp)

cplx dotc(const std::vector<cplx>& a, const std::vector<cplx>& b) {
  cplx s(0.0, 0.0);
  for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
  return s;
}

double norm2(const std::vector<cplx>& v) {
  double s = 0.0;
  for (const cplx& z : v) s += std::norm(z);
  return s;
}

void gauss_start(long dim, unsigned seed, std::vector<cplx>* q) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> g(0.0, 1.0);
  q->resize(static_cast<size_t>(dim));
  for (long i = 0; i < dim; ++i) {
    const double re = g(rng);
    const double im = g(rng);
    (*q)[static_cast<size_t>(i)] = cplx(re, im);
  }
}

template <class ApplyFn>
int lanczos(long dim, ApplyFn&& apply, int M, unsigned seed, const cplx* q0,
            std::vector<double>* alpha, std::vector<double>* beta) {
  alpha->clear();
  beta->clear();
  if (dim <= 0 || M <= 0) return 0;
  std::vector<cplx> q, qp(static_cast<size_t>(dim), cplx(0.0, 0.0)), w(static_cast<size_t>(dim));
  if (q0 != nullptr) {
    q.assign(q0, q0 + dim);
  } else {
    gauss_start(dim, seed, &q);
  }
  const double n0 = std::sqrt(norm2(q));
  if (n0 < 1e-18) {
    q[0] = cplx(1.0, 0.0);
  } else {
    const cplx s(1.0 / n0, 0.0);
    for (cplx& z : q) z *= s;
  }
  double bprev = 0.0;
  int used = 0;
  for (int k = 0; k < M; ++k) {
    apply(q.data(), w.data());
    const cplx z = dotc(q, w);
    for (long i = 0; i < dim; ++i) w[i] -= z * q[i];
    if (k > 0)
      for (long i = 0; i < dim; ++i) w[i] -= bprev * qp[i];
    const double nw2 = norm2(w);
    alpha->push_back(z.real());
    if (nw2 < 1e-24) {
      used = k + 1;
      break;
    }
    const double bk = std::sqrt(nw2);
    beta->push_back(bk);
    std::swap(q, qp);
    for (long i = 0; i < dim; ++i) q[i] = w[i] / bk;
    bprev = bk;
    used = k + 1;
  }
  return used;
}

// ---------------------------------------------------------------- Jacobi (src/lanczos.cpp, src/ftlm_thermo.cpp)

// Max-pivot single-rotation Jacobi on a dense row-major n x n matrix, at most `cap` rotations,
// stopping when the largest |off-diagonal| drops below `tol`.  V (if non-null) accumulates the
// rotations starting from the identity.
void jacobi(std::vector<double>& A, std::vector<double>* V, int n, int cap, double tol) {
  if (V) {
    V->assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) (*V)[static_cast<size_t>(i) * n + i] = 1.0;
  }
  if (n <= 1) return;
  auto at = [&](int i, int j) -> double& { return A[static_cast<size_t>(i) * n + j]; };
  for (int it = 0; it < cap; ++it) {
    int p = 0, q = 1;
    double big = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        const double v = std::abs(at(i, j));
        if (v > big) { big = v; p = i; q = j; }
      }
    if (big < tol) break;
    const double app = at(p, p), aqq = at(q, q), apq = at(p, q);
    const double phi = 0.5 * std::atan2(2.0 * apq, aqq - app);
    const double c = std::cos(phi), s = std::sin(phi);
    for (int k = 0; k < n; ++k) {
      if (k == p || k == q) continue;
      const double apk = at(p, k), aqk = at(q, k);
      const double rp = c * apk - s * aqk;
      const double rq = c * aqk + s * apk;
      at(p, k) = rp;
      at(k, p) = rp;
      at(q, k) = rq;
      at(k, q) = rq;
    }
    const double npp = c * c * app - 2.0 * s * c * apq + s * s * aqq;
    const double nqq = s * s * app + 2.0 * s * c * apq + c * c * aqq;
    at(p, q) = 0.0;
    at(q, p) = 0.0;
    at(p, p) = npp;
    at(q, q) = nqq;
    if (V) {
      for (int i = 0; i < n; ++i) {
        double& vp = (*V)[static_cast<size_t>(i) * n + p];
        double& vq = (*V)[static_cast<size_t>(i) * n + q];
        const double a = vp, bq = vq;
        vp = c * a - s * bq;
        vq = s * a + c * bq;
      }
    }
  }
}

std::vector<double> tridiag_dense(const double* alpha, int n, const double* beta, int nb) {
  std::vector<double> T(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) T[static_cast<size_t>(i) * n + i] = alpha[i];
  for (int i = 0; i < n - 1 && i < nb; ++i) {
    T[static_cast<size_t>(i) * n + i + 1] = beta[i];
    T[static_cast<size_t>(i + 1) * n + i] = beta[i];
  }
  return T;
}

double quadrature(const double* alpha, int n, const double* beta, int nb, double bt) {
  if (n <= 0) return -std::numeric_limits<double>::infinity();
  if (n == 1) return -bt * alpha[0];
  std::vector<double> T = tridiag_dense(alpha, n, beta, nb), V;
  jacobi(T, &V, n, std::max(3000, 120 * n), 1e-14);
  double mx = -std::numeric_limits<double>::infinity();
  for (int k = 0; k < n; ++k) {
    const double w = V[k] * V[k];
    if (w <= 0.0) continue;
    mx = std::max(mx, -bt * T[static_cast<size_t>(k) * n + k] + std::log(w));
  }
  if (!std::isfinite(mx)) return -std::numeric_limits<double>::infinity();
  double s = 0.0;
  for (int k = 0; k < n; ++k) {
    const double w = V[k] * V[k];
    if (w <= 0.0) continue;
    s += std::exp(-bt * T[static_cast<size_t>(k) * n + k] + std::log(w) - mx);
  }
  return mx + std::log(s);
}

double log_mean_exp_plus_lndim(const std::vector<double>& lb, long dim) {
  if (lb.empty()) return -std::numeric_limits<double>::infinity();
  double mx = lb[0];
  for (double x : lb) mx = std::max(mx, x);
  double s = 0.0;
  for (double x : lb) s += std::exp(x - mx);
  return std::log(static_cast<double>(dim)) + (mx + std::log(s / static_cast<double>(lb.size())));
}

// ---------------------------------------------------------------- translations (include/ftlm/symmetry)

uint32_t tr_bits(uint32_t bits, int lx, int ly, int dx, int dy) {
  uint32_t out = 0;
  for (int s = 0; s < lx * ly; ++s) {
    if (!((bits >> s) & 1u)) continue;
    int nx = (s % lx + dx) % lx, ny = (s / lx + dy) % ly;
    if (nx < 0) nx += lx;
    if (ny < 0) ny += ly;
    out |= 1u << (ny * lx + nx);
  }
  return out;
}

uint32_t pack(uint32_t u, uint32_t d) { return (u << 16) | d; }

double tr_sign(uint32_t u, uint32_t d, int lx, int ly, int dx, int dy) {
  // include/ftlm/symmetry/fermionic_translation.hpp:32-54: parity of inversions of the translated
  // mode list (sites ascending, up before down per site).
  std::vector<int> modes;
  for (int s = 0; s < lx * ly; ++s) {
    int nx = (s % lx + dx) % lx, ny = (s / lx + dy) % ly;
    if (nx < 0) nx += lx;
    if (ny < 0) ny += ly;
    const int ts = ny * lx + nx;
    if ((u >> s) & 1u) modes.push_back(2 * ts);
    if ((d >> s) & 1u) modes.push_back(2 * ts + 1);
  }
  int inv = 0;
  for (size_t i = 0; i < modes.size(); ++i)
    for (size_t j = i + 1; j < modes.size(); ++j)
      if (modes[i] > modes[j]) ++inv;
  return (inv & 1) ? -1.0 : 1.0;
}

cplx bloch(int kx, int ky, int lx, int ly, int dx, int dy) {
  // include/ftlm/symmetry/momentum_sector.hpp:57-63
  const double pi = std::acos(-1.0);
  const double th = -2.0 * pi *
                    (static_cast<double>(kx * dx) / static_cast<double>(lx) +
                     static_cast<double>(ky * dy) / static_cast<double>(ly));
  return cplx(std::cos(th), std::sin(th));
}

uint32_t canonical(uint32_t u, uint32_t d, int lx, int ly) {
  uint32_t best = 0xFFFFFFFFu;
  for (int dy = 0; dy < ly; ++dy)
    for (int dx = 0; dx < lx; ++dx)
      best = std::min(best, pack(tr_bits(u, lx, ly, dx, dy), tr_bits(d, lx, ly, dx, dy)));
  return best;
}

struct OrbitEntry {
  uint32_t rep;  // packed (up << 16) | dn
  int orbit_size;
  int stab_size;
  cplx A;
  double absA;
};

struct OrbitBasis {
  int lx = 0, ly = 0, kx = 0, ky = 0;
  std::vector<OrbitEntry> e;
  std::unordered_map<uint32_t, int> row;
};

void orbit_build(const Basis& fb, int lx, int ly, int kx, int ky, double tol, OrbitBasis* ob) {
  ob->lx = lx;
  ob->ly = ly;
  ob->kx = kx;
  ob->ky = ky;
  ob->e.clear();
  ob->row.clear();
  const long dim = fb.dim();
  for (long i = 0; i < dim; ++i) {
    const uint32_t u = static_cast<uint32_t>(fb.up(i)), d = static_cast<uint32_t>(fb.dn(i));
    const uint32_t p = pack(u, d);
    if (canonical(u, d, lx, ly) != p) continue;  // keep the orbit's minimal-pack member
    // distinct images -> orbit size; stabiliser in (dy outer, dx inner) order
    std::vector<uint32_t> imgs;
    cplx A(0.0, 0.0);
    int stab = 0;
    for (int dy = 0; dy < ly; ++dy)
      for (int dx = 0; dx < lx; ++dx) {
        const uint32_t q = pack(tr_bits(u, lx, ly, dx, dy), tr_bits(d, lx, ly, dx, dy));
        imgs.push_back(q);
        if (q == p) {
          ++stab;
          A += cplx(tr_sign(u, d, lx, ly, dx, dy), 0.0) * bloch(kx, ky, lx, ly, dx, dy);
        }
      }
    std::sort(imgs.begin(), imgs.end());
    imgs.erase(std::unique(imgs.begin(), imgs.end()), imgs.end());
    const double a = std::abs(A);
    if (a <= tol) continue;
    ob->e.push_back(OrbitEntry{p, static_cast<int>(imgs.size()), stab, A, a});
  }
  std::sort(ob->e.begin(), ob->e.end(),
            [](const OrbitEntry& a, const OrbitEntry& b) { return a.rep < b.rep; });
  for (size_t r = 0; r < ob->e.size(); ++r) ob->row[ob->e[r].rep] = static_cast<int>(r);
}

bool orbit_map(const OrbitBasis& ob, uint32_t u, uint32_t d, int* ri, int* dx, int* dy, double* eta) {
  const uint32_t rp = canonical(u, d, ob.lx, ob.ly);
  auto it = ob.row.find(rp);
  if (it == ob.row.end()) return false;
  const uint32_t ru = rp >> 16, rd = rp & 0xFFFFu;
  for (int tx = 0; tx < ob.lx; ++tx)
    for (int ty = 0; ty < ob.ly; ++ty) {
      if (tr_bits(ru, ob.lx, ob.ly, tx, ty) == u && tr_bits(rd, ob.lx, ob.ly, tx, ty) == d) {
        *ri = it->second;
        *dx = tx;
        *dy = ty;
        *eta = tr_sign(ru, rd, ob.lx, ob.ly, tx, ty);
        return true;
      }
    }
  return false;
}

void orbit_apply(const Model& m, const OrbitBasis& ob, const std::vector<Hop>& hops,
                 const std::vector<Pair>& pairs, const cplx* x, cplx* y) {
  const int dk = static_cast<int>(ob.e.size());
  std::fill(y, y + dk, cplx(0.0, 0.0));
  for (int b = 0; b < dk; ++b) {
    const cplx xb = x[b];
    if (std::norm(xb) < 1e-28) continue;
    const OrbitEntry& eb = ob.e[static_cast<size_t>(b)];
    const uint64_t u = eb.rep >> 16, d = eb.rep & 0xFFFFu;
    cplx dg(static_cast<double>(__builtin_popcountll(u & d)) * m.U, 0.0);
    for (const Pair& pr : pairs) dg += m.V * static_cast<double>(occ(u, d, pr.i) * occ(u, d, pr.j));
    y[b] += dg * xb;
    for (const Hop& h : hops) {
      uint64_t u2 = u, d2 = d;
      const uint64_t src = h.spin == 0 ? u : d;
      if (!((src >> h.from) & 1) || ((src >> h.to) & 1)) continue;
      const uint64_t flip = (uint64_t{1} << h.from) ^ (uint64_t{1} << h.to);
      if (h.spin == 0) u2 ^= flip; else d2 ^= flip;
      const int nb = between(2 * h.from + h.spin, 2 * h.to + h.spin, u, d);
      const cplx hsb = ((nb & 1) ? -1.0 : 1.0) * h.coeff;
      int a = -1, dx = 0, dy = 0;
      double eta = 0.0;
      if (!orbit_map(ob, static_cast<uint32_t>(u2), static_cast<uint32_t>(d2), &a, &dx, &dy, &eta)) continue;
      const double nrm = std::sqrt(ob.e[static_cast<size_t>(a)].absA / eb.absA);
      const cplx chi = bloch(ob.kx, ob.ky, ob.lx, ob.ly, dx, dy);
      y[a] += hsb * cplx(eta * nrm, 0.0) * chi * xb;
    }
  }
}

// ---------------------------------------------------------------- converged tridiagonal eigensolver

// Implicit symmetric QL with Wilkinson shift on (d, e); z holds the first row of the eigenvector
// matrix (the only part the Gauss quadrature needs).  Textbook algorithm (Golub & Van Loan 8.3).
int tridiag_ql(int n, double* d, double* e, double* z0) {
  for (int i = 0; i < n; ++i) z0[i] = (i == 0) ? 1.0 : 0.0;
  if (n <= 1) return 0;
  std::vector<double> ee(static_cast<size_t>(n), 0.0);
  for (int i = 0; i < n - 1; ++i) ee[i] = e[i];
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
        if (std::abs(ee[m]) <= 1e-300 + std::numeric_limits<double>::epsilon() * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) return -1;
      double g = (d[l + 1] - d[l]) / (2.0 * ee[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + ee[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      for (; i >= l; --i) {
        double f = s * ee[i];
        const double b = c * ee[i];
        r = std::hypot(f, g);
        ee[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          ee[m] = 0.0;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        f = z0[i + 1];
        z0[i + 1] = s * z0[i] + c * f;
        z0[i] = c * z0[i] - s * f;
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= p;
      ee[l] = g;
      ee[m] = 0.0;
    }
  }
  return 0;
}

}  // namespace

// ====================================================================== C ABI for ctypes
extern "C" {

// params8 = {Lx, Ly, t, tp, U, V, phi_x, phi_y}
int orc_geometry(const double* params8, int* from, int* to, int* spin, double* re, double* im,
                 int hop_cap, int* n_hops, int* pi, int* pj, int pair_cap, int* n_pairs) {
  const Model m = model_from(params8);
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  *n_hops = static_cast<int>(hops.size());
  *n_pairs = static_cast<int>(pairs.size());
  if (static_cast<int>(hops.size()) > hop_cap || static_cast<int>(pairs.size()) > pair_cap) return 1;
  for (size_t k = 0; k < hops.size(); ++k) {
    from[k] = hops[k].from;
    to[k] = hops[k].to;
    spin[k] = hops[k].spin;
    re[k] = hops[k].coeff.real();
    im[k] = hops[k].coeff.imag();
  }
  for (size_t k = 0; k < pairs.size(); ++k) {
    pi[k] = pairs[k].i;
    pj[k] = pairs[k].j;
  }
  return 0;
}

long orc_basis_dim(int n, int nu, int nd) {
  Basis b;
  if (!make_basis(n, nu, nd, &b, false)) return -1;
  return b.dim();
}

int orc_basis_masks(int n, int nu, int nd, uint64_t* up, uint64_t* dn) {
  Basis b;
  if (!make_basis(n, nu, nd, &b, false)) return 1;
  for (long i = 0; i < b.dim(); ++i) {
    up[i] = b.up(i);
    dn[i] = b.dn(i);
  }
  return 0;
}

int orc_index_of(int n, int nu, int nd, const uint64_t* up, const uint64_t* dn, long cnt, long* out) {
  Basis b;
  if (!make_basis(n, nu, nd, &b, true)) return 1;
  for (long i = 0; i < cnt; ++i) out[i] = b.index_of(up[i], dn[i]);
  return 0;
}

int orc_apply(const double* params8, int nu, int nd, const double* x, double* y) {
  const Model m = model_from(params8);
  Basis b;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &b, true)) return 1;
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  apply_full(m, b, hops, pairs, reinterpret_cast<const cplx*>(x), reinterpret_cast<cplx*>(y));
  return 0;
}

// Push-form H*x restricted to columns [c0, c1) with the closed-form rank in place of the hash map
// (identical results; the hash map would need ~10 GB at 4x4 (8,8)).  y is accumulated, not zeroed.
// Used for the bounded CPU baseline sample.
int orc_apply_cols(const double* params8, int nu, int nd, const double* x, double* y, long c0, long c1) {
  const Model m = model_from(params8);
  Basis b;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &b, false)) return 1;
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  Ranker ru, rd;
  ru.init(b.n, nu);
  rd.init(b.n, nd);
  const long D = static_cast<long>(b.dns.size());
  if (c0 < 0 || c1 > b.dim() || c0 > c1) return 2;
  apply_cols(m, b, hops, pairs, reinterpret_cast<const cplx*>(x), reinterpret_cast<cplx*>(y), c0, c1,
             [&](uint64_t u, uint64_t d) { return ru.rank(u) * D + rd.rank(d); });
  return 0;
}

// Selected rows of y = H x with the reference's push semantics: for each requested row r, every
// column c that can reach r (r itself and the states one listed hop away from r) is pushed
// through the full hop loop of src/hubbard_hamiltonian.cpp:58-107 and only the contributions that
// land on r are kept, in column order.  Lets a test check rows of a 4x4 (8,8) product without
// the ~60 min full push.  Ranks use the closed form (no hash map).
int orc_apply_rows(const double* params8, int nu, int nd, const double* x, long nrows, const long* rows, double* out) {
  const Model m = model_from(params8);
  Basis b;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &b, false)) return 1;
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  Ranker ru, rd;
  ru.init(b.n, nu);
  rd.init(b.n, nd);
  const long D = static_cast<long>(b.dns.size());
  const cplx* xc = reinterpret_cast<const cplx*>(x);
  for (long q = 0; q < nrows; ++q) {
    const long r = rows[q];
    if (r < 0 || r >= b.dim()) return 2;
    const uint64_t ur = b.up(r), dr = b.dn(r);
    std::vector<long> cols{r};
    for (const Hop& h : hops) {  // states one hop away (the list holds both directions)
      const uint64_t s = h.spin == 0 ? ur : dr;
      if (!((s >> h.from) & 1) || ((s >> h.to) & 1)) continue;
      const uint64_t flip = (uint64_t{1} << h.from) ^ (uint64_t{1} << h.to);
      const uint64_t u2 = h.spin == 0 ? ur ^ flip : ur, d2 = h.spin == 1 ? dr ^ flip : dr;
      cols.push_back(ru.rank(u2) * D + rd.rank(d2));
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    cplx acc(0.0, 0.0);
    std::vector<cplx> ytmp(1);
    for (long c : cols) {
      // push column c, keep what lands on row r
      const uint64_t u = b.up(c), d = b.dn(c);
      const cplx vc = xc[c];
      if (std::norm(vc) < 1e-28) continue;
      if (c == r) {
        cplx dg(static_cast<double>(__builtin_popcountll(u & d)) * m.U, 0.0);
        for (const Pair& pr : pairs) dg += m.V * static_cast<double>(occ(u, d, pr.i) * occ(u, d, pr.j));
        acc += dg * vc;
      }
      for (const Hop& h : hops) {
        const uint64_t src = h.spin == 0 ? u : d;
        if (!((src >> h.from) & 1) || ((src >> h.to) & 1)) continue;
        const uint64_t flip = (uint64_t{1} << h.from) ^ (uint64_t{1} << h.to);
        const uint64_t u2 = h.spin == 0 ? u ^ flip : u, d2 = h.spin == 1 ? d ^ flip : d;
        if (ru.rank(u2) * D + rd.rank(d2) != r) continue;
        const int nb = between(2 * h.from + h.spin, 2 * h.to + h.spin, u, d);
        acc += (((nb & 1) ? -1.0 : 1.0) * h.coeff) * vc;
      }
    }
    out[2 * q] = acc.real();
    out[2 * q + 1] = acc.imag();
  }
  return 0;
}

int orc_rank_closed_form(int n, int k, const uint64_t* masks, long cnt, long* out) {
  Ranker r;
  r.init(n, k);
  for (long i = 0; i < cnt; ++i) out[i] = r.rank(masks[i]);
  return 0;
}

int orc_gauss_start(long dim, unsigned seed, double* q) {
  std::vector<cplx> v;
  gauss_start(dim, seed, &v);
  std::memcpy(q, v.data(), sizeof(cplx) * static_cast<size_t>(dim));
  return 0;
}

// Returns steps used; alpha/beta must hold M entries; *nb receives len(beta).
int orc_lanczos(const double* params8, int nu, int nd, int M, unsigned seed, const double* q0,
                double* alpha, double* beta, int* nb) {
  const Model m = model_from(params8);
  Basis b;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &b, true)) return -1;
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  std::vector<double> a, be;
  const int used = lanczos(
      b.dim(), [&](const cplx* x, cplx* y) { apply_full(m, b, hops, pairs, x, y); }, M, seed,
      reinterpret_cast<const cplx*>(q0), &a, &be);
  std::copy(a.begin(), a.end(), alpha);
  std::copy(be.begin(), be.end(), beta);
  *nb = static_cast<int>(be.size());
  return used;
}

int orc_tridiagonal_extrema(const double* alpha, int n, const double* beta, int nb, double* lo, double* hi) {
  if (n <= 0) { *lo = 0.0; *hi = 0.0; return 0; }
  std::vector<double> T = tridiag_dense(alpha, n, beta, nb);
  jacobi(T, nullptr, n, 120, 1e-15);
  double a = std::numeric_limits<double>::infinity(), c = -a;
  for (int i = 0; i < n; ++i) {
    a = std::min(a, T[static_cast<size_t>(i) * n + i]);
    c = std::max(c, T[static_cast<size_t>(i) * n + i]);
  }
  *lo = a;
  *hi = c;
  return 0;
}

// lanczos_extrema (src/lanczos.cpp:211-231): returns steps used.
int orc_lanczos_extrema(const double* params8, int nu, int nd, int M, unsigned seed, double* lo, double* hi) {
  std::vector<double> a(static_cast<size_t>(std::max(M, 1))), be(static_cast<size_t>(std::max(M, 1)));
  int nb = 0;
  const int used = orc_lanczos(params8, nu, nd, M, seed, nullptr, a.data(), be.data(), &nb);
  *lo = 0.0;
  *hi = 0.0;
  if (used <= 0) return used;
  orc_tridiagonal_extrema(a.data(), used, be.data(), used >= 2 ? used - 1 : 0, lo, hi);
  return used;
}

double orc_quadrature(const double* alpha, int n, const double* beta, int nb, double bt) {
  return quadrature(alpha, n, beta, nb, bt);
}

// Ritz values and first-component weights of the capped Jacobi used by the quadrature.
int orc_quadrature_ritz(const double* alpha, int n, const double* beta, int nb, double* lam, double* w) {
  if (n <= 0) return 0;
  if (n == 1) { lam[0] = alpha[0]; w[0] = 1.0; return 1; }
  std::vector<double> T = tridiag_dense(alpha, n, beta, nb), V;
  jacobi(T, &V, n, std::max(3000, 120 * n), 1e-14);
  for (int k = 0; k < n; ++k) {
    lam[k] = T[static_cast<size_t>(k) * n + k];
    w[k] = V[k] * V[k];
  }
  return n;
}

int orc_tridiag_eig_ql(const double* alpha, int n, const double* beta, double* lam, double* w) {
  std::vector<double> d(alpha, alpha + n), e(static_cast<size_t>(std::max(n, 1)), 0.0), z(static_cast<size_t>(n));
  for (int i = 0; i < n - 1; ++i) e[i] = beta[i];
  const int rc = tridiag_ql(n, d.data(), e.data(), z.data());
  for (int k = 0; k < n; ++k) {
    lam[k] = d[k];
    w[k] = z[k] * z[k];
  }
  return rc;
}

// ftlm_log_partition_complex (src/ftlm_thermo.cpp:213-256); real=1 gives the _real variant with
// the apply_real wrapper of benchmarks/bench_ftlm_nmu_rect.cpp:151-159 (real part of complex apply).
double orc_ftlm_log_partition(const double* params8, int nu, int nd, double bt, int R, int M,
                              unsigned seed, int real) {
  const Model m = model_from(params8);
  Basis b;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &b, true)) return std::nan("");
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  const long dim = b.dim();
  if (dim <= 0) return -std::numeric_limits<double>::infinity();
  std::vector<cplx> tx(static_cast<size_t>(dim)), ty(static_cast<size_t>(dim));
  // real variant: H(v_re + i v_im) = Re(H v_re) + i Re(H v_im)
  auto apply_c = [&](const cplx* x, cplx* y) {
    if (!real) {
      apply_full(m, b, hops, pairs, x, y);
      return;
    }
    std::vector<double> yr(static_cast<size_t>(dim)), yi(static_cast<size_t>(dim));
    for (long i = 0; i < dim; ++i) tx[i] = cplx(x[i].real(), 0.0);
    apply_full(m, b, hops, pairs, tx.data(), ty.data());
    for (long i = 0; i < dim; ++i) yr[i] = ty[i].real();
    for (long i = 0; i < dim; ++i) tx[i] = cplx(x[i].imag(), 0.0);
    apply_full(m, b, hops, pairs, tx.data(), ty.data());
    for (long i = 0; i < dim; ++i) yi[i] = ty[i].real();
    for (long i = 0; i < dim; ++i) y[i] = cplx(yr[i], yi[i]);
  };
  if (dim == 1) {
    cplx x1[1] = {cplx(1.0, 0.0)}, y1[1] = {cplx(0.0, 0.0)};
    apply_c(x1, y1);
    return -bt * y1[0].real();
  }
  const int RR = std::max(1, R), MM = std::max(1, M);
  std::vector<double> lb;
  for (int r = 0; r < RR; ++r) {
    std::vector<double> a, be;
    const unsigned sr = seed + static_cast<unsigned>(r) * 100003u;
    const int used = lanczos(dim, apply_c, MM, sr, nullptr, &a, &be);
    if (used <= 0) continue;
    lb.push_back(quadrature(a.data(), static_cast<int>(a.size()), be.data(), static_cast<int>(be.size()), bt));
  }
  return log_mean_exp_plus_lndim(lb, dim);
}

// Grand-canonical <N>/n_sites from per-sector lnZ (benchmarks/bench_ftlm_nmu_rect.cpp:186-219).
void orc_gc_density(const double* lnZ, const int* N, int nsec, double bt, const double* mu, int nmu,
                    int n_sites, double* n_out) {
  for (int k = 0; k < nmu; ++k) {
    double mx = -std::numeric_limits<double>::infinity();
    for (int i = 0; i < nsec; ++i) {
      if (!std::isfinite(lnZ[i])) continue;
      mx = std::max(mx, lnZ[i] + bt * mu[k] * static_cast<double>(N[i]));
    }
    if (!std::isfinite(mx)) { n_out[k] = 0.0; continue; }
    double sw = 0.0, snw = 0.0;
    for (int i = 0; i < nsec; ++i) {
      if (!std::isfinite(lnZ[i])) continue;
      const double w = std::exp(lnZ[i] + bt * mu[k] * static_cast<double>(N[i]) - mx);
      sw += w;
      snw += static_cast<double>(N[i]) * w;
    }
    n_out[k] = (sw > 0.0) ? (snw / sw) / static_cast<double>(n_sites) : 0.0;
  }
}

// ---- orbit k-blocks.  reps (packed), orbit sizes, stabiliser sizes, A (re,im) for basis rows.
int orc_orbit_basis(int lx, int ly, int nu, int nd, int kx, int ky, double tol, uint32_t* reps,
                    int* orbit_size, int* stab_size, double* A, long cap) {
  Basis fb;
  if (!make_basis(lx * ly, nu, nd, &fb, false)) return -1;
  if (lx * ly > 16) return -2;
  OrbitBasis ob;
  orbit_build(fb, lx, ly, kx, ky, tol, &ob);
  const long dk = static_cast<long>(ob.e.size());
  if (reps != nullptr && dk <= cap) {
    for (long r = 0; r < dk; ++r) {
      reps[r] = ob.e[r].rep;
      orbit_size[r] = ob.e[r].orbit_size;
      stab_size[r] = ob.e[r].stab_size;
      A[2 * r] = ob.e[r].A.real();
      A[2 * r + 1] = ob.e[r].A.imag();
    }
  }
  return static_cast<int>(dk);
}

// map_raw_to_rep over a batch of packed states: out4 = (row, dx, dy, eta) or row = -1.
int orc_orbit_map(int lx, int ly, int nu, int nd, int kx, int ky, double tol, const uint32_t* states,
                  long cnt, int* row, int* dx, int* dy, double* eta) {
  Basis fb;
  if (!make_basis(lx * ly, nu, nd, &fb, false)) return -1;
  OrbitBasis ob;
  orbit_build(fb, lx, ly, kx, ky, tol, &ob);
  for (long i = 0; i < cnt; ++i) {
    int r = -1, a = 0, c = 0;
    double e = 0.0;
    if (!orbit_map(ob, states[i] >> 16, states[i] & 0xFFFFu, &r, &a, &c, &e)) r = -1;
    row[i] = r;
    dx[i] = a;
    dy[i] = c;
    eta[i] = e;
  }
  return 0;
}

int orc_orbit_apply(const double* params8, int nu, int nd, int kx, int ky, double tol, const double* x,
                    double* y) {
  const Model m = model_from(params8);
  Basis fb;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &fb, false)) return -1;
  OrbitBasis ob;
  orbit_build(fb, m.Lx, m.Ly, kx, ky, tol, &ob);
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  orbit_apply(m, ob, hops, pairs, reinterpret_cast<const cplx*>(x), reinterpret_cast<cplx*>(y));
  return 0;
}

// Selected rows of y = H_K x for the orbit block (apply_orbit_k_block, src/orbit_k_block_basis.cpp:151-220)
// without building the sector map: the caller gives the block's sorted representatives (reps,
// nreps).  For row a every push that lands on it is enumerated: each distinct member s' of a's
// orbit and each listed hop h that can produce s' name the column state s = h^-1(s'); when s is
// a representative of the block (canonical, |A| > tol, found in reps) the reference's push term
//   jw(s, h) * coeff(h) * eta(rep_a, R) * sqrt(|A_a| / |A_b|) * translation_bloch_phase(K, R)
// is added, with R the first translation (dx outer, dy inner) taking rep_a to s'
// (map_raw_to_rep, src/orbit_k_block_basis.cpp:127-149) and the jw sign evaluated on the column
// state s as the reference does.  The diagonal term is added first.  A representative that the
// oracle finds in the block but the list lacks (or the reverse) returns 3.
int orc_orbit_apply_rows(const double* params8, int nu, int nd, int kx, int ky, double tol, const uint32_t* reps,
                         long nreps, const double* x, long nrows, const long* rows, double* out) {
  const Model m = model_from(params8);
  const int lx = m.Lx, ly = m.Ly, n = lx * ly;
  if (n > 16 || nu < 0 || nd < 0 || nu > n || nd > n) return 1;
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  const cplx* xc = reinterpret_cast<const cplx*>(x);
  // |A(K)| of a canonical state (orbit_build's stabiliser sum, dy outer / dx inner)
  auto absA = [&](uint32_t p) {
    const uint32_t u = p >> 16, d = p & 0xFFFFu;
    cplx A(0.0, 0.0);
    for (int dy = 0; dy < ly; ++dy)
      for (int dx = 0; dx < lx; ++dx)
        if (pack(tr_bits(u, lx, ly, dx, dy), tr_bits(d, lx, ly, dx, dy)) == p)
          A += cplx(tr_sign(u, d, lx, ly, dx, dy), 0.0) * bloch(kx, ky, lx, ly, dx, dy);
    return std::abs(A);
  };
  for (long q = 0; q < nrows; ++q) {
    const long a = rows[q];
    if (a < 0 || a >= nreps) return 2;
    const uint32_t pa = reps[a];
    const uint32_t ua = pa >> 16, da = pa & 0xFFFFu;
    if (canonical(ua, da, lx, ly) != pa) return 3;
    const double Aa = absA(pa);
    if (Aa <= tol) return 3;
    cplx acc(0.0, 0.0);
    {
      cplx dg(static_cast<double>(__builtin_popcount(ua & da)) * m.U, 0.0);
      for (const Pair& pr : pairs) dg += m.V * static_cast<double>(occ(ua, da, pr.i) * occ(ua, da, pr.j));
      acc += dg * xc[a];
    }
    std::vector<uint32_t> members;
    for (int dx = 0; dx < lx; ++dx)
      for (int dy = 0; dy < ly; ++dy) members.push_back(pack(tr_bits(ua, lx, ly, dx, dy), tr_bits(da, lx, ly, dx, dy)));
    std::sort(members.begin(), members.end());
    members.erase(std::unique(members.begin(), members.end()), members.end());
    for (uint32_t sp : members) {
      const uint32_t u2 = sp >> 16, d2 = sp & 0xFFFFu;
      // R: first (dx outer, dy inner) translation with T_R rep_a = s'
      int rx = -1, ry = -1;
      for (int tx = 0; tx < lx && rx < 0; ++tx)
        for (int ty = 0; ty < ly; ++ty)
          if (tr_bits(ua, lx, ly, tx, ty) == u2 && tr_bits(da, lx, ly, tx, ty) == d2) {
            rx = tx;
            ry = ty;
            break;
          }
      const double eta = tr_sign(ua, da, lx, ly, rx, ry);
      const cplx chi = bloch(kx, ky, lx, ly, rx, ry);
      for (const Hop& h : hops) {
        const uint32_t tgt = h.spin == 0 ? u2 : d2;  // s' must hold `to` and lack `from`
        if (h.from == h.to || !((tgt >> h.to) & 1u) || ((tgt >> h.from) & 1u)) continue;
        const uint32_t flip = (1u << h.from) ^ (1u << h.to);
        const uint32_t us = h.spin == 0 ? u2 ^ flip : u2, ds = h.spin == 1 ? d2 ^ flip : d2;
        const uint32_t ps = pack(us, ds);
        if (canonical(us, ds, lx, ly) != ps) continue;  // columns are representatives
        const uint32_t* it = std::lower_bound(reps, reps + nreps, ps);
        const bool listed = it != reps + nreps && *it == ps;
        const double Ab = absA(ps);
        if ((Ab > tol) != listed) return 3;
        if (!listed) continue;
        const long b = static_cast<long>(it - reps);
        const cplx xb = xc[b];
        if (std::norm(xb) < 1e-28) continue;
        const int nb = between(2 * h.from + h.spin, 2 * h.to + h.spin, us, ds);
        const cplx hsb = ((nb & 1) ? -1.0 : 1.0) * h.coeff;
        acc += hsb * cplx(eta * std::sqrt(Aa / Ab), 0.0) * chi * xb;
      }
    }
    out[2 * q] = acc.real();
    out[2 * q + 1] = acc.imag();
  }
  return 0;
}

// Full-sector lift Phi_K v of a block vector v over the block's sorted representatives: column a
// of Phi is fill_phi_orbit_bloch_from_seed(seed = rep_a) (include/ftlm/symmetry/orbit_bloch_phi.hpp:35-51),
//   sum over all (ex, ey) of c * eta(rep_a, e) * conj(translation_bloch_phase(K, e)) at T_e rep_a,
// c = 1 / sqrt(|G| |Stab_a|) (momentum_sector.hpp:74-77).  Columns of distinct orbits touch
// disjoint states, so the sum is split over `threads` workers.  out (complex, C(n,nu) C(n,nd))
// is overwritten.
int orc_orbit_lift(int lx, int ly, int nu, int nd, int kx, int ky, const uint32_t* reps, long nreps, const double* v,
                   double* out, int threads) {
  const int n = lx * ly;
  if (n > 16 || nu < 0 || nd < 0 || nu > n || nd > n) return 1;
  Ranker ru, rd;
  ru.init(n, nu);
  rd.init(n, nd);
  const long Dd = rd.binom(n, nd), dim = ru.binom(n, nu) * Dd;
  const cplx* vc = reinterpret_cast<const cplx*>(v);
  cplx* oc = reinterpret_cast<cplx*>(out);
  std::fill(oc, oc + dim, cplx(0.0, 0.0));
  const double G = static_cast<double>(lx * ly);
  auto work = [&](long a0, long a1) {
    for (long a = a0; a < a1; ++a) {
      const uint32_t u = reps[a] >> 16, d = reps[a] & 0xFFFFu;
      int stab = 0;
      for (int ey = 0; ey < ly; ++ey)
        for (int ex = 0; ex < lx; ++ex)
          stab += tr_bits(u, lx, ly, ex, ey) == u && tr_bits(d, lx, ly, ex, ey) == d;
      const double c = 1.0 / std::sqrt(G * static_cast<double>(stab));
      for (int ey = 0; ey < ly; ++ey)
        for (int ex = 0; ex < lx; ++ex) {
          const uint32_t u2 = tr_bits(u, lx, ly, ex, ey), d2 = tr_bits(d, lx, ly, ex, ey);
          const long idx = ru.rank(u2) * Dd + rd.rank(d2);
          const double eta = tr_sign(u, d, lx, ly, ex, ey);
          oc[idx] += c * eta * std::conj(bloch(kx, ky, lx, ly, ex, ey)) * vc[a];
        }
    }
  };
  const int nt = std::max(1, std::min(threads, 64));
  std::vector<std::thread> pool;
  const long per = (nreps + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const long a0 = std::min(nreps, t * per), a1 = std::min(nreps, a0 + per);
    if (a0 < a1) pool.emplace_back(work, a0, a1);
  }
  for (auto& th : pool) th.join();
  return 0;
}

double orc_translation_sign(uint32_t u, uint32_t d, int lx, int ly, int dx, int dy) {
  return tr_sign(u, d, lx, ly, dx, dy);
}

uint32_t orc_canonical(uint32_t u, uint32_t d, int lx, int ly) { return canonical(u, d, lx, ly); }

// Packed image of (u, d) under translation (dx, dy) (occup_bits_translate.hpp:10-25).
uint32_t orc_translate(uint32_t u, uint32_t d, int lx, int ly, int dx, int dy) {
  return pack(tr_bits(u, lx, ly, dx, dy), tr_bits(d, lx, ly, dx, dy));
}

// Lanczos on the orbit block with ftlm_log_partition_complex semantics (one beta).
double orc_orbit_ftlm_log_partition(const double* params8, int nu, int nd, int kx, int ky, double bt,
                                    int R, int M, unsigned seed) {
  const Model m = model_from(params8);
  Basis fb;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &fb, false)) return std::nan("");
  OrbitBasis ob;
  orbit_build(fb, m.Lx, m.Ly, kx, ky, 1e-12, &ob);
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  const long dk = static_cast<long>(ob.e.size());
  if (dk <= 0) return -std::numeric_limits<double>::infinity();
  auto ap = [&](const cplx* x, cplx* y) { orbit_apply(m, ob, hops, pairs, x, y); };
  if (dk == 1) {
    cplx x1[1] = {cplx(1.0, 0.0)}, y1[1];
    ap(x1, y1);
    return -bt * y1[0].real();
  }
  std::vector<double> lb;
  for (int r = 0; r < std::max(1, R); ++r) {
    std::vector<double> a, be;
    const int used = lanczos(dk, ap, std::max(1, M), seed + static_cast<unsigned>(r) * 100003u, nullptr, &a, &be);
    if (used <= 0) continue;
    lb.push_back(quadrature(a.data(), static_cast<int>(a.size()), be.data(), static_cast<int>(be.size()), bt));
  }
  return log_mean_exp_plus_lndim(lb, dk);
}

// exact_log_partition_small_complex (benchmarks/bench_ftlm_nmu_4x4_k.cpp:249-336) applied to a
// caller-supplied dense Hermitian matrix H (row-major complex, d x d): 2d x 2d real embedding,
// Jacobi with max(5000,150*2d) rotations at tol 1e-13, log-sum-exp, + ln(1/2).
double orc_exact_log_partition(const double* H, int d, double bt) {
  if (d <= 0) return -std::numeric_limits<double>::infinity();
  if (d == 1) return -bt * H[0];
  const int m2 = 2 * d;
  std::vector<double> R(static_cast<size_t>(m2) * m2, 0.0);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      const double re = H[2 * (static_cast<size_t>(i) * d + j)], im = H[2 * (static_cast<size_t>(i) * d + j) + 1];
      R[static_cast<size_t>(i) * m2 + j] = re;
      R[static_cast<size_t>(i) * m2 + j + d] = -im;
      R[static_cast<size_t>(i + d) * m2 + j] = im;
      R[static_cast<size_t>(i + d) * m2 + j + d] = re;
    }
  jacobi(R, nullptr, m2, std::max(5000, 150 * m2), 1e-13);
  double mx = -std::numeric_limits<double>::infinity();
  for (int i = 0; i < m2; ++i) mx = std::max(mx, -bt * R[static_cast<size_t>(i) * m2 + i]);
  double s = 0.0;
  for (int i = 0; i < m2; ++i) s += std::exp(-bt * R[static_cast<size_t>(i) * m2 + i] - mx);
  return std::log(0.5) + mx + std::log(s);
}

// Dense sector matrix from unit-vector applies (benchmarks/bench_2x2.cpp:157-178), complex
// row-major d x d, and its 1600-rotation Jacobi extrema (bench_2x2.cpp:28-93) when real.
int orc_dense_hamiltonian(const double* params8, int nu, int nd, double* H) {
  const Model m = model_from(params8);
  Basis b;
  if (!make_basis(m.Lx * m.Ly, nu, nd, &b, true)) return 1;
  std::vector<Hop> hops;
  std::vector<Pair> pairs;
  geometry(m, &hops, &pairs);
  const long dim = b.dim();
  std::vector<cplx> e(static_cast<size_t>(dim)), y(static_cast<size_t>(dim));
  for (long j = 0; j < dim; ++j) {
    std::fill(e.begin(), e.end(), cplx(0.0, 0.0));
    e[j] = cplx(1.0, 0.0);
    apply_full(m, b, hops, pairs, e.data(), y.data());
    for (long i = 0; i < dim; ++i) {
      H[2 * (i * dim + j)] = y[i].real();
      H[2 * (i * dim + j) + 1] = y[i].imag();
    }
  }
  return 0;
}

int orc_dense_extrema(const double* Hreal, int n, int cap, double* lo, double* hi) {
  std::vector<double> A(Hreal, Hreal + static_cast<size_t>(n) * n);
  jacobi(A, nullptr, n, cap, 1e-15);
  double a = std::numeric_limits<double>::infinity(), c = -a;
  for (int i = 0; i < n; ++i) {
    a = std::min(a, A[static_cast<size_t>(i) * n + i]);
    c = std::max(c, A[static_cast<size_t>(i) * n + i]);
  }
  *lo = a;
  *hi = c;
  return 0;
}

}  // extern "C"
