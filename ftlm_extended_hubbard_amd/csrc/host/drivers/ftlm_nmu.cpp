// ftlm_nmu.cpp -- grand-canonical n(mu) at fixed beta on the device: the sector / momentum-block
// orchestration of benchmarks/bench_ftlm_nmu_rect.cpp (built as ftlm_nmu_rect) and
// benchmarks/bench_ftlm_nmu_4x4_k.cpp (built as ftlm_nmu_rect_k, FTLM_NMU_K=1), over the C ABI.
//
// Kept as the reference has them (they are the drivers' interface): the --key=value flags and
// their defaults, the per-sector seed (every sector uses --seed) and per-block seed
// seed + idx*257 + ky*17 + kx (bench_ftlm_nmu_4x4_k.cpp:1077), the ED cutoff for small blocks
// (:587, 1072), the grand-canonical bookkeeping (bench_ftlm_nmu_rect.cpp:186-219) and the TSV /
// METRIC output (:181-184, 232-244).  Every sector or block runs on the device:
//   nonmom   ftlm_log_partition_real of every (N_up, N_dn) sector, all sectors in one
//            ftlm_hip_ftlm_ritz call (one persistent launch when every sector is small);
//   k-gram   (the default, as in the reference) every momentum block as HubbardMomentumBlock: the
//            Gram-basis block applied through the full sector (ftlm_hip_kgram_create); FTLM
//            (complex) above the ED cutoff, device dense ED below;
//   k-orbit  the blocks of sectors with 0 < N_up, N_dn < N and at most 512 states as
//            orbit-representative operators (apply_orbit_k_block), every other block as a Gram
//            block, as the reference gates its orbit-direct path (:797-807, 1003-1019);
//   ed       every block a Gram block, by device dense ED.
// The Gram blocks keep the reference's subspace (orbits with trivial bosonic and fermionic
// stabiliser characters); their orthonormal basis is not the reference's zheev basis, so ED blocks
// agree exactly and FTLM blocks in distribution.  Device options beyond the reference:
// --start=gauss|phase, --quad=capped|converged, --dtype=c128|c64 (defaults reproduce the reference),
// --orbit-all=1 (k-orbit: orbit blocks in every sector) and --path=auto|per-op|stack (the FTLM
// scheduling: the orbit blocks of a sector run as one ftlm_hip_ftlm_ritz_seeded call, which stacks
// them under auto where that is faster and under stack always; random-phase starts only).
#include <sys/resource.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "driver_host.hpp"
#include "ftlm_hip.h"

#ifndef FTLM_NMU_K
#define FTLM_NMU_K 0
#endif

namespace {

using ftlm_driver::arg_double;
using ftlm_driver::arg_int;
using ftlm_driver::find_arg;
using ftlm_driver::has_arg;
using ftlm_driver::logsumexp2;
using ftlm_driver::write_nmu;
using clock_type = std::chrono::high_resolution_clock;

double peak_rss_bytes() {
  struct rusage ru {};
  if (getrusage(RUSAGE_SELF, &ru) != 0) return 0.0;
  return static_cast<double>(ru.ru_maxrss) * 1024.0;
}

[[noreturn]] void fail(const char* what) {
  std::cerr << "ftlm_nmu: " << what << ": " << ftlm_hip_last_error() << "\n";
  std::exit(3);
}
void ok(int rc, const char* what) {
  if (rc != FTLM_OK) fail(what);
}

struct Model {
  int Lx = 4, Ly = 2;
  ftlm_params_t p{};
  std::vector<ftlm_hop_t> hops;
  std::vector<ftlm_pair_t> pairs;
};

Model make_model(int argc, char** argv, int Lx, int Ly) {
  Model m;
  m.Lx = Lx;
  m.Ly = Ly;
  m.p.Lx = Lx;
  m.p.Ly = Ly;
  m.p.t = arg_double(argc, argv, "t", 1.0);
  m.p.tp = arg_double(argc, argv, "tp", -0.35);
  m.p.U = arg_double(argc, argv, "U", 5.75);
  m.p.V = arg_double(argc, argv, "V", 0.9);
  m.p.phi_x = 0.0;
  m.p.phi_y = 0.0;
  int nh = 0, np = 0;
  ftlm_hip_build_geometry(&m.p, nullptr, 0, &nh, nullptr, 0, &np);  // sizes (FTLM_ERANGE expected)
  m.hops.resize(static_cast<size_t>(std::max(nh, 1)));
  m.pairs.resize(static_cast<size_t>(std::max(np, 1)));
  ok(ftlm_hip_build_geometry(&m.p, m.hops.data(), nh, &nh, m.pairs.data(), np, &np), "build_hubbard_geometry");
  m.hops.resize(static_cast<size_t>(nh));
  m.pairs.resize(static_cast<size_t>(np));
  return m;
}

// Device options: defaults give the reference's estimator (mt19937 Gaussian starts, capped
// Jacobi quadrature, complex128 vectors).
ftlm_ftlm_opts_t make_opts(int argc, char** argv, int R, int M, unsigned seed, int real_variant) {
  ftlm_ftlm_opts_t o{};
  o.R = R;
  o.M = M;
  o.seed = seed;
  const char* st = find_arg(argc, argv, "start");
  const char* qd = find_arg(argc, argv, "quad");
  const char* dt = find_arg(argc, argv, "dtype");
  o.start = (st && std::strcmp(st, "phase") == 0) ? FTLM_START_PHASE : FTLM_START_GAUSS_MT;
  o.quad = (qd && std::strcmp(qd, "converged") == 0) ? FTLM_QUAD_CONVERGED : FTLM_QUAD_REFERENCE;
  o.dtype = (dt && std::strcmp(dt, "c64") == 0) ? FTLM_C64 : FTLM_C128;
  o.clamp_M_to_dim = 0;
  o.real_variant = real_variant;
  const char* pa = find_arg(argc, argv, "path");  // not in the reference: the device scheduling
  o.path = !pa                             ? FTLM_PATH_AUTO
           : std::strcmp(pa, "per-op") == 0 ? FTLM_PATH_PER_OP
           : std::strcmp(pa, "stack") == 0  ? FTLM_PATH_STACK
                                            : FTLM_PATH_AUTO;
  return o;
}

int sector_index(int nu, int nd, int n_sites) { return nu * (n_sites + 1) + nd; }

struct Common {
  int Lx, Ly;
  double beta, mu_min, mu_max;
  int n_mu, n_rand, lz_steps;
  unsigned seed;
  const char* out_path;
  int no_monitor;
};

Common parse_common(int argc, char** argv, int Ly_dflt, int nrand_dflt, int lz_dflt) {
  Common c;
  c.Lx = arg_int(argc, argv, "Lx", 4);
  c.Ly = arg_int(argc, argv, "Ly", Ly_dflt);
  c.beta = arg_double(argc, argv, "beta", 20.0);
  c.mu_min = arg_double(argc, argv, "mu-min", -5.0);
  c.mu_max = arg_double(argc, argv, "mu-max", 25.0);
  c.n_mu = std::max(2, arg_int(argc, argv, "n-mu", 221));
  c.n_rand = arg_int(argc, argv, "ftlm-random", nrand_dflt);
  c.lz_steps = has_arg(argc, argv, "lanczos-steps") ? arg_int(argc, argv, "lanczos-steps", lz_dflt) : lz_dflt;
  c.seed = static_cast<unsigned>(arg_int(argc, argv, "seed", 7));
  c.out_path = find_arg(argc, argv, "out");
  c.no_monitor = arg_int(argc, argv, "no-monitor", 0);
  return c;
}

// Non-momentum solver: ftlm_log_partition_real of every sector (bench_ftlm_nmu_rect.cpp:92-246).
int run_nonmom(int argc, char** argv, const char* metric_kind) {
  const Common c = parse_common(argc, argv, 2, 12, 72);
  const auto t_wall0 = clock_type::now();
  const Model m = make_model(argc, argv, c.Lx, c.Ly);
  const int n_sites = c.Lx * c.Ly;
  const int nsec = (n_sites + 1) * (n_sites + 1);
  std::vector<double> logZ(static_cast<size_t>(nsec), -std::numeric_limits<double>::infinity());
  std::vector<int> Nelec(static_cast<size_t>(nsec), 0);
  std::cout << "=== FTLM grand-canonical n(mu)  " << c.Lx << "x" << c.Ly << "  beta=" << c.beta << " ===\n";
  std::cout << "t=" << m.p.t << " tp=" << m.p.tp << " U=" << m.p.U << " V=" << m.p.V << "\n";
  std::cout << "ftlm_random=" << c.n_rand << " lanczos_steps=" << c.lz_steps << "  sectors=" << nsec << "\n";
  if (!c.no_monitor) std::cout << "[monitor] wall time + peak RSS reported at exit (disable with --no-monitor=1)\n";

  const auto t0 = clock_type::now();
  // every non-empty sector as a device operator, then one FTLM call over all of them
  std::vector<ftlm_op_t*> ops;
  std::vector<int> idx_of_op, N_of_op;
  std::vector<int64_t> dims;
  for (int nu = 0; nu <= n_sites; ++nu)
    for (int nd = 0; nd <= n_sites; ++nd) {
      const int idx = sector_index(nu, nd, n_sites);
      Nelec[static_cast<size_t>(idx)] = nu + nd;
      ftlm_op_t* op = nullptr;
      ok(ftlm_hip_sector_create(n_sites, nu, nd, m.hops.data(), static_cast<int>(m.hops.size()), m.pairs.data(),
                                static_cast<int>(m.pairs.size()), m.p.U, m.p.V, &op),
         "sector_create");
      int64_t dim = 0;
      ok(ftlm_hip_op_dim(op, &dim, nullptr, nullptr), "op_dim");
      if (dim <= 0) {
        ftlm_hip_op_destroy(op);
        continue;
      }
      ops.push_back(op);
      idx_of_op.push_back(idx);
      N_of_op.push_back(nu + nd);
      dims.push_back(dim);
    }
  const ftlm_ftlm_opts_t o = make_opts(argc, argv, c.n_rand, c.lz_steps, c.seed, 1);
  const int nops = static_cast<int>(ops.size());
  const size_t per = static_cast<size_t>(c.n_rand) * c.lz_steps;
  std::vector<double> ritz(per * nops), weight(per * nops), lz(static_cast<size_t>(nops));
  std::vector<int> used(static_cast<size_t>(c.n_rand) * nops);
  ok(ftlm_hip_ftlm_ritz(ops.data(), nops, &o, ritz.data(), weight.data(), used.data(), nullptr, nullptr, nullptr),
     "ftlm_ritz");
  ok(ftlm_hip_ftlm_thermo(ritz.data(), weight.data(), used.data(), nops, c.n_rand, c.lz_steps, dims.data(),
                          N_of_op.data(), &c.beta, 1, nullptr, 0, lz.data(), nullptr, nullptr, nullptr, nullptr),
     "ftlm_thermo");
  for (int i = 0; i < nops; ++i) {
    logZ[static_cast<size_t>(idx_of_op[i])] = lz[i];
    const int nu = idx_of_op[i] / (n_sites + 1), nd = idx_of_op[i] % (n_sites + 1);
    std::cerr << "sector (" << nu << "," << nd << ") dim=" << dims[i] << " logZ=" << lz[i] << "\n";
    ftlm_hip_op_destroy(ops[i]);
  }
  const double t_sectors = std::chrono::duration<double>(clock_type::now() - t0).count();
  std::cout << "[timing] all sectors logZ: " << t_sectors << " s\n";

  std::ostream* out = &std::cout;
  std::ofstream file;
  if (c.out_path) {
    file.open(c.out_path);
    if (!file) {
      std::cerr << "ftlm_nmu_rect: cannot open " << c.out_path << "\n";
      return 2;
    }
    out = &file;
  }
  *out << "# FTLM grand canonical  Lx=" << c.Lx << " Ly=" << c.Ly << "  beta=" << c.beta << "  t=" << m.p.t
       << " tp=" << m.p.tp << " U=" << m.p.U << " V=" << m.p.V << "\n";
  *out << "# ftlm_random=" << c.n_rand << " lanczos_steps=" << c.lz_steps << "  sector_time_s=" << t_sectors << "\n";
  write_nmu(*out, logZ, Nelec, n_sites, c.beta, c.mu_min, c.mu_max, c.n_mu);
  if (c.out_path) std::cout << "wrote " << c.out_path << "\n";
  const double t_wall = std::chrono::duration<double>(clock_type::now() - t_wall0).count();
  const double peak_b = peak_rss_bytes();
  const double peak_mib = peak_b / (1024.0 * 1024.0);
  if (!c.no_monitor) {
    std::cout << "[monitor] wall_time_s=" << t_wall << "  wall_sector_logZ_s=" << t_sectors
              << "  peak_rss_mib=" << peak_mib << "  peak_rss_bytes=" << peak_b << "\n";
    if (c.out_path && file.is_open()) {
      *out << "# wall_time_s=" << t_wall << "  peak_rss_bytes=" << peak_b << "  peak_rss_mib=" << peak_mib << "\n";
      file.flush();
    }
  }
  std::cout << "METRIC kind=" << metric_kind << " Lx=" << c.Lx << " Ly=" << c.Ly << " beta=" << c.beta
            << " wall_time_s=" << t_wall << " wall_sector_logZ_s=" << t_sectors << " peak_rss_bytes=" << peak_b
            << " peak_rss_mib=" << peak_mib << " n_mu=" << c.n_mu;
  if (c.no_monitor) std::cout << " no_monitor=1";
  std::cout << "\n";
  return 0;
}

#if FTLM_NMU_K
// Momentum-block solvers (bench_ftlm_nmu_4x4_k.cpp:546-1262): k-gram, k-orbit and ed.
enum class KSolver { kGram, kOrbit, kEd };
constexpr int kOrbitDirectMaxDim = 512;  // kOrbitMfMaxDim, bench_ftlm_nmu_4x4_k.cpp:616

const char* k_label(KSolver s) { return s == KSolver::kGram ? "k-gram" : s == KSolver::kOrbit ? "k-orbit" : "ed"; }

int run_k(int argc, char** argv, KSolver solver, bool solver_set) {
  const bool ed_only = solver == KSolver::kEd;
  const int Lx = arg_int(argc, argv, "Lx", 4), Ly = arg_int(argc, argv, "Ly", 3);
  if (Lx <= 0 || Ly <= 0 || Lx * Ly > 16) {
    std::cerr << "ftlm_nmu_rect_k: requires 1 <= Lx*Ly <= 16\n";
    return 2;
  }
  const int n_sites = Lx * Ly;
  // momentum-Lanczos depth default: max(20, 2 * 72 / N)
  const Common c = parse_common(argc, argv, 3, 8, std::max(20, (2 * 72) / std::max(1, n_sites)));
  const bool log_k_dims = arg_int(argc, argv, "no-log-k-dims", 0) == 0;
  int ed_cutoff = arg_int(argc, argv, "ed-cutoff", 64);
  if (ed_only && !has_arg(argc, argv, "ed-cutoff")) ed_cutoff = std::numeric_limits<int>::max() / 4;
  const auto t_wall0 = clock_type::now();
  const Model m = make_model(argc, argv, Lx, Ly);
  const int nsec = (n_sites + 1) * (n_sites + 1);
  std::vector<double> logZ(static_cast<size_t>(nsec), -std::numeric_limits<double>::infinity());
  std::vector<int> Nelec(static_cast<size_t>(nsec), 0);
  std::cout << "=== FTLM grand-canonical n(mu) " << Lx << "x" << Ly << " with momentum blocks ===\n";
  if (solver_set) {  // bench_ftlm_nmu_4x4_k.cpp:704-710
    std::cout << "[solver] " << k_label(solver);
    if (ed_only) std::cout << "  ed_cutoff=" << ed_cutoff;
    std::cout << "\n";
  }
  std::cout << "beta=" << c.beta << "  t=" << m.p.t << " tp=" << m.p.tp << " U=" << m.p.U << " V=" << m.p.V << "\n";
  std::cout << "ftlm_random=" << c.n_rand << " lanczos_steps=" << c.lz_steps << "  sectors=" << nsec << "\n";
  if (!c.no_monitor) std::cout << "[monitor] wall time + peak RSS reported at exit (disable with --no-monitor=1)\n";

  // --orbit-all=1 (not in the reference): orbit blocks in every sector of the k-orbit solver.  They
  // keep every state and skip the full-sector applies of the Gram blocks.
  const bool orbit_all = arg_int(argc, argv, "orbit-all", 0) != 0;
  if (solver == KSolver::kOrbit)  // bench_ftlm_nmu_4x4_k.cpp:736-740
    std::cout << "[kblock-prototype] orbit-direct: OrbitKBlockBasis + apply_orbit_k_block in orbit coords (no Phi "
                 "Gram, no T bridge); dim_full<="
              << kOrbitDirectMaxDim << "\n";
  const auto t0 = clock_type::now();
  int64_t blocks_ed = 0, blocks_ftlm = 0, orbit_direct_used = 0, orbit_direct_zero = 0;
  double t_create = 0.0, t_ed = 0.0, t_ftlm = 0.0;  // per-phase wall time, reported on stderr
  auto since = [](clock_type::time_point a) { return std::chrono::duration<double>(clock_type::now() - a).count(); };
  for (int nu = 0; nu <= n_sites; ++nu)
    for (int nd = 0; nd <= n_sites; ++nd) {
      const int idx = sector_index(nu, nd, n_sites);
      Nelec[static_cast<size_t>(idx)] = nu + nd;
      double logZ_sector = -std::numeric_limits<double>::infinity();
      int64_t dim_k_total = 0;
      int n_ed = 0, n_ftlm = 0;
      // the sector's first block is kept until the sector is done: later blocks share its orbit
      // graph; every other block is released as soon as its logZ is known.  Gram blocks apply
      // through the full sector operator, built once per sector.
      ftlm_op_t* first = nullptr;
      ftlm_op_t* sector_op = nullptr;
      int64_t dim_full = 0;
      {
        ftlm_op_t* probe = nullptr;  // FockBasis dim (sector_create validates and enumerates)
        ok(ftlm_hip_sector_create(n_sites, nu, nd, m.hops.data(), static_cast<int>(m.hops.size()),
                                  m.pairs.data(), static_cast<int>(m.pairs.size()), m.p.U, m.p.V, &probe),
           "sector_create");
        ok(ftlm_hip_op_dim(probe, &dim_full, nullptr, nullptr), "op_dim");
        sector_op = probe;
      }
      if (dim_full <= 0) {  // bench_ftlm_nmu_4x4_k.cpp:752-755
        ftlm_hip_op_destroy(sector_op);
        continue;
      }
      // orbit coordinates only for the k-orbit solver's interior sectors of <= 512 states
      // (bench_ftlm_nmu_4x4_k.cpp:800-807); HubbardMomentumBlock everywhere else (:1003-1019)
      const bool orbit_direct =
          solver == KSolver::kOrbit &&
          (orbit_all || (nu > 0 && nd > 0 && nu < n_sites && nd < n_sites && dim_full <= kOrbitDirectMaxDim));
      if (orbit_direct) {
        ftlm_hip_op_destroy(sector_op);
        sector_op = nullptr;
      }
      // logZ of every K (ED blocks and Gram blocks at once; the orbit blocks' FTLM runs as one call
      // per sector after the K loop, so that the blocks can share momentum stacks), summed in K order
      std::vector<double> lz_of_k(static_cast<size_t>(Lx * Ly), -std::numeric_limits<double>::infinity());
      std::vector<ftlm_op_t*> pend_op, held;
      std::vector<unsigned> pend_seed;
      std::vector<int> pend_k;
      for (int ky = 0; ky < Ly; ++ky)
        for (int kx = 0; kx < Lx; ++kx) {
          ftlm_op_t* op = nullptr;
          const auto tc = clock_type::now();
          ok(ftlm_hip_kblock_create(Lx, Ly, nu, nd, kx, ky, 1e-12, m.hops.data(), static_cast<int>(m.hops.size()),
                                    m.pairs.data(), static_cast<int>(m.pairs.size()), m.p.U, m.p.V, first, &op),
             "kblock_create");
          t_create += since(tc);
          bool keep = first == nullptr;
          if (keep) first = op;
          struct Release {
            ftlm_op_t* op;
            const bool* keep;
            ~Release() {
              if (!*keep) ftlm_hip_op_destroy(op);
            }
          } release{op, &keep};
          ftlm_op_t* gram = nullptr;
          if (!orbit_direct) ok(ftlm_hip_kgram_create(sector_op, op, &gram), "kgram_create");
          struct ReleaseGram {
            ftlm_op_t* op;
            ~ReleaseGram() {
              if (op) ftlm_hip_op_destroy(op);
            }
          } release_gram{gram};
          ftlm_op_t* const blk = gram ? gram : op;
          int64_t dk = 0;
          ok(ftlm_hip_op_dim(blk, &dk, nullptr, nullptr), "op_dim");
          dim_k_total += dk;
          if (orbit_direct) ++(dk > 0 ? orbit_direct_used : orbit_direct_zero);
          if (log_k_dims)
            std::cerr << "  k=(" << kx << "," << ky << ") dim=" << dk << (gram ? "" : " [orbit-direct]") << "\n";
          if (dk <= 0) continue;
          double lz_k = -std::numeric_limits<double>::infinity();
          const auto ts = clock_type::now();
          const unsigned seed_k = c.seed + static_cast<unsigned>(idx * 257 + ky * 17 + kx);
          if (dk <= ed_cutoff) {
            ++n_ed;
            ok(ftlm_hip_exact_log_partition(blk, &c.beta, 1, &lz_k, nullptr), "exact_log_partition");
            t_ed += since(ts);
          } else if (gram == nullptr) {  // an orbit block: kept for the sector's batched FTLM
            ++n_ftlm;
            pend_op.push_back(op);
            pend_seed.push_back(seed_k);
            pend_k.push_back(ky * Lx + kx);
            if (!keep) held.push_back(op);
            keep = true;
            continue;
          } else {
            ++n_ftlm;
            const ftlm_ftlm_opts_t o = make_opts(argc, argv, c.n_rand, c.lz_steps, seed_k, 0);
            ok(ftlm_hip_ftlm_log_partition(blk, &c.beta, 1, &o, &lz_k), "ftlm_log_partition");
            t_ftlm += since(ts);
          }
          lz_of_k[static_cast<size_t>(ky * Lx + kx)] = lz_k;
        }
      if (!pend_op.empty()) {
        // ftlm_log_partition of every pending block, as one call (bench_ftlm_nmu_4x4_k.cpp:1077-1091)
        const auto ts = clock_type::now();
        const int np = static_cast<int>(pend_op.size());
        const ftlm_ftlm_opts_t o = make_opts(argc, argv, c.n_rand, c.lz_steps, pend_seed[0], 0);
        const size_t per = static_cast<size_t>(c.n_rand) * c.lz_steps;
        std::vector<double> ritz(per * np), weight(per * np), lz(static_cast<size_t>(np));
        std::vector<int> used(static_cast<size_t>(c.n_rand) * np);
        std::vector<int64_t> dims(static_cast<size_t>(np));
        for (int i = 0; i < np; ++i) ok(ftlm_hip_op_dim(pend_op[i], &dims[i], nullptr, nullptr), "op_dim");
        ok(ftlm_hip_ftlm_ritz_seeded(pend_op.data(), np, &o, pend_seed.data(), ritz.data(), weight.data(),
                                     used.data(), nullptr, nullptr, nullptr),
           "ftlm_ritz_seeded");
        ok(ftlm_hip_ftlm_thermo(ritz.data(), weight.data(), used.data(), np, c.n_rand, c.lz_steps, dims.data(),
                                nullptr, &c.beta, 1, nullptr, 0, lz.data(), nullptr, nullptr, nullptr, nullptr),
           "ftlm_thermo");
        for (int i = 0; i < np; ++i)  // a single state: -beta H00 (ftlm_log_partition)
          lz_of_k[static_cast<size_t>(pend_k[i])] = dims[i] == 1 ? -c.beta * ritz[per * i] : lz[i];
        t_ftlm += since(ts);
      }
      for (ftlm_op_t* h : held) ftlm_hip_op_destroy(h);
      for (double lz_k : lz_of_k)
        if (lz_k > -std::numeric_limits<double>::infinity()) logZ_sector = logsumexp2(logZ_sector, lz_k);
      if (first) ftlm_hip_op_destroy(first);
      if (sector_op) ftlm_hip_op_destroy(sector_op);
      blocks_ed += n_ed;
      blocks_ftlm += n_ftlm;
      logZ[static_cast<size_t>(idx)] = logZ_sector;
      std::cerr << "sector (" << nu << "," << nd << ") dim_full=" << dim_full << " dim_k_total=" << dim_k_total
                << " ed_blocks=" << n_ed
                << " ftlm_blocks=" << n_ftlm << " logZ=" << logZ_sector << "\n";
    }
  const double t_sectors = std::chrono::duration<double>(clock_type::now() - t0).count();
  std::cout << "[timing] all sectors logZ: " << t_sectors << " s\n";
  std::cerr << "[timing] kblock_create " << t_create << " s, ed " << t_ed << " s (" << blocks_ed << " blocks), ftlm "
            << t_ftlm << " s (" << blocks_ftlm << " blocks)\n";

  std::ostream* out = &std::cout;
  std::ofstream file;
  if (c.out_path) {
    file.open(c.out_path);
    if (!file) {
      std::cerr << "ftlm_nmu_rect_k: cannot open " << c.out_path << "\n";
      return 2;
    }
    out = &file;
  }
  *out << "# FTLM grand canonical (rect k-block) Lx=" << Lx << " Ly=" << Ly << " beta=" << c.beta << " t=" << m.p.t
       << " tp=" << m.p.tp << " U=" << m.p.U << " V=" << m.p.V << "\n";
  *out << "# ftlm_random=" << c.n_rand << " lanczos_steps=" << c.lz_steps << " sector_time_s=" << t_sectors << "\n";
  write_nmu(*out, logZ, Nelec, n_sites, c.beta, c.mu_min, c.mu_max, c.n_mu);
  if (c.out_path) std::cout << "wrote " << c.out_path << "\n";
  const double t_wall = std::chrono::duration<double>(clock_type::now() - t_wall0).count();
  const double peak_b = peak_rss_bytes();
  const double peak_mib = peak_b / (1024.0 * 1024.0);
  if (!c.no_monitor) {
    std::cout << "[monitor] wall_time_s=" << t_wall << "  wall_sector_logZ_s=" << t_sectors
              << "  peak_rss_mib=" << peak_mib << "  peak_rss_bytes=" << peak_b << "\n";
    if (c.out_path && file.is_open()) {
      *out << "# wall_time_s=" << t_wall << " peak_rss_bytes=" << peak_b << " peak_rss_mib=" << peak_mib << "\n";
      file.flush();
    }
  }
  std::cout << "METRIC kind=FTLM_nmu_rect_k solver=" << (solver_set ? k_label(solver) : "default")
            << " wall_time_s=" << t_wall << " wall_sector_logZ_s=" << t_sectors << " peak_rss_bytes=" << peak_b
            << " peak_rss_mib=" << peak_mib << " n_mu=" << c.n_mu << " gram_build_no_reuse=0 kblock_prototype="
            << (solver == KSolver::kOrbit ? "orbit-direct" : "none");
  if (solver == KSolver::kOrbit)
    std::cout << " orbit_direct_blocks_used=" << orbit_direct_used << " orbit_direct_skipped_zero=" << orbit_direct_zero;
  std::cout << " ed_blocks=" << blocks_ed << " ftlm_blocks=" << blocks_ftlm << "\n";
  return 0;
}
#endif

}  // namespace

int main(int argc, char** argv) {
#if FTLM_NMU_K
  // solver names as parse_solver_name accepts them (bench_ftlm_nmu_4x4_k.cpp:340-362)
  const char* solver = find_arg(argc, argv, "solver");
  auto is = [&](std::initializer_list<const char*> names) {
    for (const char* n : names)
      if (std::strcmp(solver, n) == 0) return true;
    return false;
  };
  if (solver == nullptr) return run_k(argc, argv, KSolver::kGram, false);
  if (is({"nonmom", "nomom", "non-momentum"})) return run_nonmom(argc, argv, "FTLM_nmu_rect solver=nonmom");
  if (is({"ed", "exact"})) return run_k(argc, argv, KSolver::kEd, true);
  if (is({"k-orbit", "orbit", "orbit-direct"})) return run_k(argc, argv, KSolver::kOrbit, true);
  if (is({"k-gram", "k_block_gram", "gram", "production"})) return run_k(argc, argv, KSolver::kGram, true);
  std::cerr << "ftlm_nmu_rect_k: unknown --solver=" << solver << " (supported: k-gram, k-orbit, nonmom, ed)\n";
  return 2;
#else
  return run_nonmom(argc, argv, "FTLM_nmu_rect");
#endif
}
