// ftlm_host.cpp -- the reference's C++ API for the FTLM hot path (namespace ftlm), implemented on
// the gfx950 library through its C ABI (include/ftlm_hip.h).  Built as libftlm.so next to
// libftlm_hip.so; a program written against the reference headers links this instead of the
// reference's ftlm_core.  Status codes become the reference's exceptions:
//   FTLM_EINVAL / FTLM_EUNSUPPORTED -> std::invalid_argument, FTLM_ERANGE -> std::out_of_range,
//   FTLM_ENOMEM -> std::bad_alloc, FTLM_EHIP -> std::runtime_error.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "ftlm/fock_basis.hpp"
#include "ftlm/ftlm_thermo.hpp"
#include "ftlm/hip/device.hpp"
#include "ftlm/hubbard_hamiltonian.hpp"
#include "ftlm/hubbard_params.hpp"
#include "ftlm/lanczos.hpp"
#include "ftlm/lattice.hpp"
#include "ftlm/symmetry/fermionic_translation.hpp"
#include "ftlm/symmetry/orbit.hpp"
#include "ftlm/symmetry/orbit_k_block_basis.hpp"
#include "ftlm/symmetry/packed_raw_state.hpp"
#include "ftlm/version.hpp"
#include "ftlm_hip.h"

namespace ftlm {
namespace {

void check(int rc) {
  if (rc == FTLM_OK) return;
  const std::string msg = ftlm_hip_last_error();
  switch (rc) {
    case FTLM_EINVAL:
    case FTLM_EUNSUPPORTED:
      throw std::invalid_argument(msg);
    case FTLM_ERANGE:
      throw std::out_of_range(msg);
    case FTLM_ENOMEM:
      throw std::bad_alloc();
    default:
      throw std::runtime_error(msg);
  }
}

std::shared_ptr<ftlm_op> own(ftlm_op* op) {
  return std::shared_ptr<ftlm_op>(op, [](ftlm_op* p) { (void)ftlm_hip_op_destroy(p); });
}

int64_t op_dim(const ftlm_op* op) {
  int64_t d = 0;
  check(ftlm_hip_op_dim(op, &d, nullptr, nullptr));
  return d;
}

ftlm_params_t c_params(const HubbardParams& p) {
  ftlm_params_t c{};
  c.Lx = p.Lx;
  c.Ly = p.Ly;
  c.t = p.t;
  c.tp = p.tp;
  c.U = p.U;
  c.V = p.V;
  c.phi_x = p.phi_x;
  c.phi_y = p.phi_y;
  return c;
}

std::vector<ftlm_hop_t> c_hops(const std::vector<SpinfulHopping>& h) {
  std::vector<ftlm_hop_t> out(h.size());
  for (size_t i = 0; i < h.size(); ++i) out[i] = ftlm_hop_t{h[i].from, h[i].to, h[i].spin, h[i].coeff.real(), h[i].coeff.imag()};
  return out;
}

std::vector<ftlm_pair_t> c_pairs(const std::vector<NearestPair>& p) {
  std::vector<ftlm_pair_t> out(p.size());
  for (size_t i = 0; i < p.size(); ++i) out[i] = ftlm_pair_t{p[i].i, p[i].j};
  return out;
}

// Cache key of a Hamiltonian: its hop list, pair list and U, V as raw bytes.
std::string model_key(const std::vector<ftlm_hop_t>& h, const std::vector<ftlm_pair_t>& p, double U, double V) {
  std::string k;
  k.append(reinterpret_cast<const char*>(h.data()), h.size() * sizeof(ftlm_hop_t));
  k.push_back('|');
  k.append(reinterpret_cast<const char*>(p.data()), p.size() * sizeof(ftlm_pair_t));
  k.append(reinterpret_cast<const char*>(&U), sizeof U);
  k.append(reinterpret_cast<const char*>(&V), sizeof V);
  return k;
}

std::shared_ptr<ftlm_op> make_sector(int n, int nu, int nd, const std::vector<ftlm_hop_t>& h,
                                     const std::vector<ftlm_pair_t>& p, double U, double V) {
  ftlm_op* op = nullptr;
  check(ftlm_hip_sector_create(n, nu, nd, h.empty() ? nullptr : h.data(), static_cast<int>(h.size()),
                               p.empty() ? nullptr : p.data(), static_cast<int>(p.size()), U, V, &op));
  return own(op);
}

// ---- host callbacks driving the device recurrence
struct CallbackCtx {
  const std::function<void(const std::complex<double>*, std::complex<double>*)>* fc = nullptr;
  const std::function<void(const double*, double*)>* fr = nullptr;
  int dim = 0;
  std::vector<double> xr, xi, yr, yi;
  std::exception_ptr err;
};

void trampoline(const double* x, double* y, void* ctx) {
  auto* c = static_cast<CallbackCtx*>(ctx);
  try {
    if (c->fc) {
      (*c->fc)(reinterpret_cast<const std::complex<double>*>(x), reinterpret_cast<std::complex<double>*>(y));
    } else {
      // real symmetric H on a complex vector: real and imaginary parts separately
      for (int i = 0; i < c->dim; ++i) {
        c->xr[i] = x[2 * i];
        c->xi[i] = x[2 * i + 1];
      }
      (*c->fr)(c->xr.data(), c->yr.data());
      (*c->fr)(c->xi.data(), c->yi.data());
      for (int i = 0; i < c->dim; ++i) {
        y[2 * i] = c->yr[i];
        y[2 * i + 1] = c->yi[i];
      }
    }
  } catch (...) {
    c->err = std::current_exception();
    throw std::runtime_error("the apply function threw");
  }
}

// The device operator behind an apply function: its own operator when it came from
// ftlm::hip, otherwise a host-callback operator around it.
struct OpFor {
  std::shared_ptr<ftlm_op> op;
  std::unique_ptr<CallbackCtx> ctx;
  bool real_variant = false;

  void call(int rc) {
    if (rc != FTLM_OK && ctx && ctx->err) std::rethrow_exception(ctx->err);
    check(rc);
  }
};

OpFor op_for(int dim, const ApplyFn& f) {
  OpFor o;
  if (const auto* d = f.target<hip::DeviceApply>()) {
    if (op_dim(d->op.get()) != dim) throw std::invalid_argument("dim does not match the device operator");
    o.op = d->op;
    return o;
  }
  o.ctx = std::make_unique<CallbackCtx>();
  o.ctx->fc = &f;
  o.ctx->dim = dim;
  ftlm_op* op = nullptr;
  check(ftlm_hip_callback_op_create(dim, trampoline, o.ctx.get(), &op));
  o.op = own(op);
  return o;
}

OpFor op_for_real(int dim, const std::function<void(const double*, double*)>& f) {
  OpFor o;
  if (const auto* d = f.target<hip::DeviceApplyReal>()) {
    if (op_dim(d->op.get()) != dim) throw std::invalid_argument("dim does not match the device operator");
    o.op = d->op;
    o.real_variant = true;
    return o;
  }
  o.ctx = std::make_unique<CallbackCtx>();
  o.ctx->fr = &f;
  o.ctx->dim = dim;
  o.ctx->xr.resize(dim);
  o.ctx->xi.resize(dim);
  o.ctx->yr.resize(dim);
  o.ctx->yi.resize(dim);
  ftlm_op* op = nullptr;
  check(ftlm_hip_callback_op_create(dim, trampoline, o.ctx.get(), &op));
  o.op = own(op);
  return o;
}

std::mutex g_cache_mu;

}  // namespace

// ------------------------------------------------------------------ params, version, geometry
int nsites(const HubbardParams& p) {
  const ftlm_params_t c = c_params(p);
  return ftlm_hip_nsites(&c);
}

const char* version_string() { return ftlm_hip_version(); }

void build_hubbard_geometry(const HubbardParams& p, const RectLattice& lat, std::vector<SpinfulHopping>* hoppings,
                            std::vector<NearestPair>* nn_pairs) {
  ftlm_params_t c = c_params(p);
  c.Lx = lat.Lx;
  c.Ly = lat.Ly;
  const int n = std::max(lat.Lx * lat.Ly, 1);
  std::vector<ftlm_hop_t> h(static_cast<size_t>(64 * n + 64));
  std::vector<ftlm_pair_t> pr(static_cast<size_t>(8 * n + 8));
  int nh = 0, np = 0;
  check(ftlm_hip_build_geometry(&c, h.data(), static_cast<int>(h.size()), &nh, pr.data(), static_cast<int>(pr.size()),
                                &np));
  hoppings->clear();
  nn_pairs->clear();
  for (int i = 0; i < nh; ++i) {
    SpinfulHopping s;
    s.from = h[i].from;
    s.to = h[i].to;
    s.spin = h[i].spin;
    s.coeff = std::complex<double>(h[i].re, h[i].im);
    hoppings->push_back(s);
  }
  for (int i = 0; i < np; ++i) nn_pairs->push_back(NearestPair{pr[i].i, pr[i].j});
}

// ------------------------------------------------------------------ FockBasis
FockBasis::FockBasis(int n_sites, int n_up, int n_down) : n_sites_(n_sites), n_up_(n_up), n_down_(n_down), dim_(0) {
  dev_ = make_sector(n_sites, n_up, n_down, {}, {}, 0.0, 0.0);
  int64_t d = 0, du = 0, dd = 0;
  check(ftlm_hip_op_dim(dev_.get(), &d, &du, &dd));
  if (d > std::numeric_limits<int>::max()) throw std::invalid_argument("FockBasis: dimension exceeds int");
  dim_ = static_cast<int>(d);
  up_str_.resize(static_cast<size_t>(du));
  dn_str_.resize(static_cast<size_t>(dd));
  check(ftlm_hip_sector_strings(dev_.get(), 0, up_str_.data()));
  check(ftlm_hip_sector_strings(dev_.get(), 1, dn_str_.data()));
  for (size_t i = 0; i < up_str_.size(); ++i) up_rank_.emplace(up_str_[i], static_cast<int>(i));
  for (size_t i = 0; i < dn_str_.size(); ++i) dn_rank_.emplace(dn_str_[i], static_cast<int>(i));
}

uint64_t FockBasis::up_mask(int idx) const {
  if (idx < 0 || idx >= dim_) throw std::out_of_range("FockBasis::up_mask: index out of range");
  return up_str_[static_cast<size_t>(idx / static_cast<int>(dn_str_.size()))];
}

uint64_t FockBasis::down_mask(int idx) const {
  if (idx < 0 || idx >= dim_) throw std::out_of_range("FockBasis::down_mask: index out of range");
  return dn_str_[static_cast<size_t>(idx % static_cast<int>(dn_str_.size()))];
}

int FockBasis::index_of(uint64_t up, uint64_t down) const {
  if ((up >> 32) || (down >> 32)) return -1;
  const auto a = up_rank_.find(static_cast<uint32_t>(up));
  const auto b = dn_rank_.find(static_cast<uint32_t>(down));
  if (a == up_rank_.end() || b == dn_rank_.end()) return -1;
  return a->second * static_cast<int>(dn_str_.size()) + b->second;
}

// ------------------------------------------------------------------ H * v
void apply_extended_hubbard(const HubbardParams& p, const FockBasis& basis, const std::vector<SpinfulHopping>& hoppings,
                            const std::vector<NearestPair>& nn_pairs, const std::complex<double>* x,
                            std::complex<double>* y) {
  static std::map<std::string, std::shared_ptr<ftlm_op>> cache;
  const auto h = c_hops(hoppings);
  const auto pr = c_pairs(nn_pairs);
  std::string key = model_key(h, pr, p.U, p.V);
  const int sec[3] = {basis.n_sites(), basis.n_up(), basis.n_down()};
  key.append(reinterpret_cast<const char*>(sec), sizeof sec);
  std::shared_ptr<ftlm_op> op;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = cache.find(key);
    if (it == cache.end()) {
      if (cache.size() > 16) cache.clear();
      it = cache.emplace(key, make_sector(sec[0], sec[1], sec[2], h, pr, p.U, p.V)).first;
    }
    op = it->second;
  }
  if (basis.dim() == 0) return;
  check(ftlm_hip_apply_host(op.get(), reinterpret_cast<const double*>(x), reinterpret_cast<double*>(y)));
}

// ------------------------------------------------------------------ Lanczos
int lanczos_tridiagonal(int dim, const ApplyFn& apply_h, int max_steps, unsigned seed, std::vector<double>* alpha,
                        std::vector<double>* beta, LanczosComplexWorkspace* ws) {
  (void)ws;
  alpha->clear();
  beta->clear();
  if (dim <= 0 || max_steps <= 0) return 0;
  OpFor o = op_for(dim, apply_h);
  std::vector<double> a(static_cast<size_t>(max_steps)), b(static_cast<size_t>(max_steps));
  int nb = 0, used = 0;
  o.call(ftlm_hip_lanczos(o.op.get(), max_steps, seed, nullptr, a.data(), b.data(), &nb, &used));
  alpha->assign(a.begin(), a.begin() + used);
  beta->assign(b.begin(), b.begin() + nb);
  return used;
}

LanczosExtrema lanczos_extrema(int dim, const ApplyFn& apply_h, int max_steps, unsigned seed) {
  LanczosExtrema out{};
  if (dim <= 0 || max_steps <= 0) return out;
  OpFor o = op_for(dim, apply_h);
  double lo = 0.0, hi = 0.0;
  int used = 0;
  o.call(ftlm_hip_lanczos_extrema(o.op.get(), max_steps, seed, &lo, &hi, &used));
  out.steps_used = used;
  if (used > 0) {
    out.min_eval = lo;
    out.max_eval = hi;
  }
  return out;
}

std::pair<double, double> tridiagonal_extrema(const std::vector<double>& alpha, const std::vector<double>& beta,
                                              int steps) {
  if (steps <= 0) return {0.0, 0.0};
  if (static_cast<int>(alpha.size()) < steps) throw std::out_of_range("tridiagonal_extrema: alpha shorter than steps");
  const int nb = std::min(static_cast<int>(beta.size()), steps - 1);
  double lo = 0.0, hi = 0.0;
  check(ftlm_hip_tridiagonal_extrema(alpha.data(), steps, nb > 0 ? beta.data() : nullptr, std::max(nb, 0), &lo, &hi));
  return {lo, hi};
}

// ------------------------------------------------------------------ FTLM
namespace {

double log_partition(OpFor& o, double beta, const FtlmParams& par) {
  ftlm_ftlm_opts_t opt{};
  opt.R = std::max(1, par.n_random);
  opt.M = std::max(1, par.lanczos_steps);
  opt.seed = par.seed;
  opt.start = FTLM_START_GAUSS_MT;
  opt.dtype = FTLM_C128;
  opt.quad = FTLM_QUAD_REFERENCE;
  opt.clamp_M_to_dim = 0;
  opt.real_variant = o.real_variant ? 1 : 0;
  double out = 0.0;
  o.call(ftlm_hip_ftlm_log_partition(o.op.get(), &beta, 1, &opt, &out));
  return out;
}

}  // namespace

double ftlm_log_partition_complex(int dim,
                                  const std::function<void(const std::complex<double>* x, std::complex<double>* y)>& apply_h,
                                  double beta, const FtlmParams& par) {
  if (dim <= 0) return -std::numeric_limits<double>::infinity();
  OpFor o = op_for(dim, apply_h);
  return log_partition(o, beta, par);
}

double ftlm_log_partition_real(int dim, const std::function<void(const double* x, double* y)>& apply_real, double beta,
                               const FtlmParams& par) {
  if (dim <= 0) return -std::numeric_limits<double>::infinity();
  OpFor o = op_for_real(dim, apply_real);
  return log_partition(o, beta, par);
}

// ------------------------------------------------------------------ device operators
namespace hip {

void DeviceApply::operator()(const std::complex<double>* x, std::complex<double>* y) const {
  check(ftlm_hip_apply_host(op.get(), reinterpret_cast<const double*>(x), reinterpret_cast<double*>(y)));
}

void DeviceApplyReal::operator()(const double* x, double* y) const {
  // (Re H) x = Re(H x) for a real vector x
  const int64_t d = op_dim(op.get());
  std::vector<std::complex<double>> xc(static_cast<size_t>(d)), yc(static_cast<size_t>(d));
  for (int64_t i = 0; i < d; ++i) xc[static_cast<size_t>(i)] = std::complex<double>(x[i], 0.0);
  check(ftlm_hip_apply_host(op.get(), reinterpret_cast<const double*>(xc.data()), reinterpret_cast<double*>(yc.data())));
  for (int64_t i = 0; i < d; ++i) y[i] = yc[static_cast<size_t>(i)].real();
}

SectorOperator::SectorOperator(const HubbardParams& p, const FockBasis& basis,
                               const std::vector<SpinfulHopping>& hoppings, const std::vector<NearestPair>& nn_pairs) {
  op_ = make_sector(basis.n_sites(), basis.n_up(), basis.n_down(), c_hops(hoppings), c_pairs(nn_pairs), p.U, p.V);
  dim_ = static_cast<int>(op_dim(op_.get()));
}

KBlockOperator::KBlockOperator(const HubbardParams& p, const FockBasis& fb, symmetry::MomentumSector K,
                               const std::vector<SpinfulHopping>& hoppings, const std::vector<NearestPair>& nn_pairs,
                               double survive_tol, const KBlockOperator* share) {
  const auto h = c_hops(hoppings);
  const auto pr = c_pairs(nn_pairs);
  if (K.lx * K.ly != fb.n_sites()) throw std::invalid_argument("KBlockOperator: K.lx * K.ly != n_sites");
  ftlm_op* op = nullptr;
  check(ftlm_hip_kblock_create(K.lx, K.ly, fb.n_up(), fb.n_down(), K.kx, K.ky, survive_tol,
                               h.empty() ? nullptr : h.data(), static_cast<int>(h.size()),
                               pr.empty() ? nullptr : pr.data(), static_cast<int>(pr.size()), p.U, p.V,
                               share ? share->op_.get() : nullptr, &op));
  op_ = own(op);
  dim_ = static_cast<int>(op_dim(op_.get()));
}

GramBlockOperator::GramBlockOperator(const SectorOperator& sector, const KBlockOperator& kblock)
    : sector_(sector.shared_handle()) {
  ftlm_op* op = nullptr;
  check(ftlm_hip_kgram_create(sector_.get(), kblock.handle(), &op));
  op_ = own(op);
  dim_ = static_cast<int>(op_dim(op_.get()));
}

}  // namespace hip

// ------------------------------------------------------------------ translations and k-blocks
namespace symmetry {

std::complex<double> fermionic_stabilizer_character_A_rect(
    const RawState& rep, MomentumSector K, int lx, int ly,
    const std::vector<std::pair<std::int8_t, std::int8_t>>& stabilizer) {
  std::complex<double> A{0.0, 0.0};
  for (const auto& e : stabilizer) {
    const double eta = fermionic_translation_sign_rect(rep, lx, ly, e.first, e.second);
    A += std::complex<double>(eta, 0.0) * translation_bloch_phase(K, e.first, e.second);
  }
  return A;
}

namespace {

// Basis-only orbit graphs, one per (lx, ly, n_up, n_down), shared by every momentum.
std::shared_ptr<ftlm_op> basis_share(int lx, int ly, int nu, int nd, std::shared_ptr<ftlm_op> candidate) {
  static std::map<std::tuple<int, int, int, int>, std::weak_ptr<ftlm_op>> reg;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto& w = reg[std::make_tuple(lx, ly, nu, nd)];
  if (auto s = w.lock()) return s;
  if (candidate) w = candidate;
  return nullptr;
}

}  // namespace

bool OrbitKBlockBasis::build(const FockBasis& fb, MomentumSector Ksec, double tol) {
  if (fb.n_sites() > 16)
    throw std::invalid_argument("OrbitKBlockBasis::build: n_sites must be <= 16 (RawState bit width)");
  if (Ksec.lx * Ksec.ly != fb.n_sites()) throw std::invalid_argument("OrbitKBlockBasis::build: K.lx * K.ly != n_sites");
  K = Ksec;
  lx = K.lx;
  ly = K.ly;
  survive_tol_ = tol;
  n_up_ = fb.n_up();
  n_dn_ = fb.n_down();
  const std::shared_ptr<ftlm_op> share = basis_share(lx, ly, n_up_, n_dn_, nullptr);
  ftlm_op* op = nullptr;
  check(ftlm_hip_kblock_create(lx, ly, n_up_, n_dn_, K.kx, K.ky, tol, nullptr, 0, nullptr, 0, 0.0, 0.0, share.get(),
                               &op));
  dev_ = own(op);
  if (!share) basis_share(lx, ly, n_up_, n_dn_, dev_);
  dim_ = static_cast<int>(op_dim(dev_.get()));
  entries_.clear();
  have_entries_ = false;
  return true;
}

const std::vector<OrbitKRepEntry>& OrbitKBlockBasis::entries() const {
  if (!have_entries_) {
    entries_.clear();
    if (dev_ && dim_ > 0) {
      std::vector<uint32_t> rep(static_cast<size_t>(dim_));
      std::vector<int> osz(static_cast<size_t>(dim_)), ssz(static_cast<size_t>(dim_));
      std::vector<double> A(2 * static_cast<size_t>(dim_));
      std::vector<int64_t> oi(static_cast<size_t>(dim_));
      check(ftlm_hip_kblock_entries(dev_.get(), 0, dim_, rep.data(), osz.data(), ssz.data(), A.data()));
      check(ftlm_hip_kblock_orbit_index(dev_.get(), 0, dim_, oi.data()));
      entries_.resize(static_cast<size_t>(dim_));
      for (int r = 0; r < dim_; ++r) {
        OrbitKRepEntry& e = entries_[static_cast<size_t>(r)];
        e.rep = unpack_raw_state(rep[r]);
        e.orbit_index = static_cast<std::size_t>(oi[r]);
        e.orbit_size = static_cast<std::uint16_t>(osz[r]);
        e.stabilizer_size = static_cast<std::uint16_t>(ssz[r]);
        e.A_k = std::complex<double>(A[2 * r], A[2 * r + 1]);
        e.abs_A = std::abs(e.A_k);
      }
    }
    have_entries_ = true;
  }
  return entries_;
}

int OrbitKBlockBasis::reduced_index_of_rep(RawState rep) const {
  if (!dev_) return -1;
  const uint32_t p = pack_raw_state(rep);
  int64_t r = -1;
  check(ftlm_hip_kblock_index_of_rep(dev_.get(), &p, 1, &r));
  return static_cast<int>(r);
}

bool OrbitKBlockBasis::map_raw_to_rep(RawState s, int* reduced_index, int* dx, int* dy, double* eta) const {
  if (reduced_index == nullptr || dx == nullptr || dy == nullptr || eta == nullptr || !dev_) return false;
  const uint32_t p = pack_raw_state(s);
  int64_t r = -1;
  int tx = 0, ty = 0;
  double e = 0.0;
  check(ftlm_hip_kblock_map(dev_.get(), &p, 1, &r, &tx, &ty, &e));
  if (r < 0) return false;
  *reduced_index = static_cast<int>(r);
  *dx = tx;
  *dy = ty;
  *eta = e;
  return true;
}

void apply_orbit_k_block(const HubbardParams& p, const FockBasis& fb, const OrbitKBlockBasis& basis,
                         const std::vector<SpinfulHopping>& hoppings, const std::vector<NearestPair>& nn_pairs,
                         const std::complex<double>* x, std::complex<double>* y) {
  static std::map<std::string, std::shared_ptr<ftlm_op>> blocks;   // per (model, sector, K, tol)
  static std::map<std::string, std::weak_ptr<ftlm_op>> graphs;     // per (model, sector)
  if (basis.dim() <= 0) return;
  const auto h = c_hops(hoppings);
  const auto pr = c_pairs(nn_pairs);
  std::string mkey = model_key(h, pr, p.U, p.V);
  const int sec[4] = {basis.lx, basis.ly, fb.n_up(), fb.n_down()};
  mkey.append(reinterpret_cast<const char*>(sec), sizeof sec);
  std::string bkey = mkey;
  const int kk[2] = {basis.K.kx, basis.K.ky};
  const double tol = basis.survive_tol();
  bkey.append(reinterpret_cast<const char*>(kk), sizeof kk);
  bkey.append(reinterpret_cast<const char*>(&tol), sizeof tol);
  std::shared_ptr<ftlm_op> op;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = blocks.find(bkey);
    if (it == blocks.end()) {
      if (blocks.size() > 64) blocks.clear();
      std::shared_ptr<ftlm_op> share = graphs[mkey].lock();
      ftlm_op* raw = nullptr;
      check(ftlm_hip_kblock_create(basis.lx, basis.ly, fb.n_up(), fb.n_down(), basis.K.kx, basis.K.ky, tol,
                                   h.empty() ? nullptr : h.data(), static_cast<int>(h.size()),
                                   pr.empty() ? nullptr : pr.data(), static_cast<int>(pr.size()), p.U, p.V,
                                   share.get(), &raw));
      auto made = own(raw);
      if (!share) graphs[mkey] = made;
      it = blocks.emplace(bkey, made).first;
    }
    op = it->second;
  }
  if (op_dim(op.get()) != basis.dim()) throw std::invalid_argument("apply_orbit_k_block: basis does not match");
  check(ftlm_hip_apply_host(op.get(), reinterpret_cast<const double*>(x), reinterpret_cast<double*>(y)));
}

}  // namespace symmetry
}  // namespace ftlm
