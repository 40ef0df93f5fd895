// ftlm_bench.cpp -- device counterparts of the reference's small benchmark drivers, over the C ABI:
//   FTLM_BENCH=1  ftlm_bench_2x2        benchmarks/bench_2x2.cpp: 2x2 (2,2) apply throughput, dense
//                                       check and Lanczos extrema against the capped dense Jacobi
//   FTLM_BENCH=2  ftlm_bench_4x2_ftlm   benchmarks/bench_4x2_ftlm.cpp: stochastic Lanczos extrema
//                                       over samples (seed + s * 7919)
//   FTLM_BENCH=3  ftlm_bench_2x2_nmu    benchmarks/bench_2x2_nmu.cpp: exact grand-canonical n(mu)
// Kept as the reference has them (they are the drivers' interface): flags and defaults, the fixed
// 2x2 parameters, seeds (mt19937 99 / 123 / lanczos-seed), Jacobi caps (1600 at 1e-15, 6000 and
// 8000 at 1e-12, 12000 at 1e-14) and the printed lines.  The applies, Lanczos recurrences and
// Jacobi rotations run on the device; a dense matrix is assembled from applies to unit vectors.
#include <sys/resource.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "driver_host.hpp"
#include "ftlm_hip.h"

#ifndef FTLM_BENCH
#define FTLM_BENCH 1
#endif

namespace {

using clock_type = std::chrono::high_resolution_clock;
using cplx = std::complex<double>;

using ftlm_driver::arg_double;
using ftlm_driver::arg_int;
using ftlm_driver::find_arg;

double seconds_since(clock_type::time_point t0) {
  return std::chrono::duration<double>(clock_type::now() - t0).count();
}

[[noreturn]] void fail(const char* what) {
  std::cerr << "ftlm_bench: " << what << ": " << ftlm_hip_last_error() << "\n";
  std::exit(3);
}
void ok(int rc, const char* what) {
  if (rc != FTLM_OK) fail(what);
}

struct Model {
  ftlm_params_t p{};
  std::vector<ftlm_hop_t> hops;
  std::vector<ftlm_pair_t> pairs;
};

Model make_model(int Lx, int Ly, double t, double tp, double U, double V) {
  Model m;
  m.p = ftlm_params_t{Lx, Ly, t, tp, U, V, 0.0, 0.0};
  int nh = 0, np = 0;
  ftlm_hip_build_geometry(&m.p, nullptr, 0, &nh, nullptr, 0, &np);  // sizes (FTLM_ERANGE expected)
  m.hops.resize(static_cast<size_t>(std::max(nh, 1)));
  m.pairs.resize(static_cast<size_t>(std::max(np, 1)));
  ok(ftlm_hip_build_geometry(&m.p, m.hops.data(), nh, &nh, m.pairs.data(), np, &np), "build_hubbard_geometry");
  m.hops.resize(static_cast<size_t>(nh));
  m.pairs.resize(static_cast<size_t>(np));
  return m;
}

ftlm_op_t* make_sector(const Model& m, int nu, int nd) {
  ftlm_op_t* op = nullptr;
  ok(ftlm_hip_sector_create(m.p.Lx * m.p.Ly, nu, nd, m.hops.data(), static_cast<int>(m.hops.size()), m.pairs.data(),
                            static_cast<int>(m.pairs.size()), m.p.U, m.p.V, &op),
     "sector_create");
  return op;
}

int64_t op_dim(const ftlm_op_t* op) {
  int64_t d = 0;
  ok(ftlm_hip_op_dim(op, &d, nullptr, nullptr), "op_dim");
  return d;
}

// Real part of H, row-major, from applies to the unit vectors (build_dense_real,
// bench_2x2.cpp:160-182); returns false when H has an imaginary part above 1e-10.
bool dense_real(const ftlm_op_t* op, int dim, std::vector<double>* H) {
  H->assign(static_cast<size_t>(dim) * dim, 0.0);
  std::vector<cplx> col(static_cast<size_t>(dim)), out(static_cast<size_t>(dim));
  for (int j = 0; j < dim; ++j) {
    std::fill(col.begin(), col.end(), cplx(0.0, 0.0));
    col[static_cast<size_t>(j)] = 1.0;
    ok(ftlm_hip_apply_host(op, reinterpret_cast<const double*>(col.data()), reinterpret_cast<double*>(out.data())),
       "apply_host");
    for (int i = 0; i < dim; ++i) {
      if (std::abs(out[static_cast<size_t>(i)].imag()) > 1e-10) return false;
      (*H)[static_cast<size_t>(i) * dim + j] = out[static_cast<size_t>(i)].real();
    }
  }
  return true;
}

std::vector<double> sorted_jacobi(const std::vector<double>& A, int n, int cap, double tol) {
  if (n <= 0) return {};
  std::vector<double> d(static_cast<size_t>(n));
  ok(ftlm_hip_dense_jacobi(A.data(), n, cap, tol, d.data(), nullptr, nullptr), "dense_jacobi");
  std::sort(d.begin(), d.end());
  return d;
}

double max_rss_mib() {
  struct rusage ru {};
  if (getrusage(RUSAGE_SELF, &ru) != 0) return 0.0;
  return static_cast<double>(ru.ru_maxrss) / 1024.0;
}

#if FTLM_BENCH == 1
int run(int argc, char** argv) {
  const long long matvec_iters = arg_int(argc, argv, "matvec-iters", 80000);
  const int lanczos_steps_req = arg_int(argc, argv, "lanczos-steps", 36);
  const unsigned lanczos_seed = static_cast<unsigned>(arg_int(argc, argv, "lanczos-seed", 7));
  const char* plot_dir = find_arg(argc, argv, "write-plot-data");
  const Model m = make_model(2, 2, 1.0, -0.35, 5.75, 0.9);
  const int n = 4, n_up = n / 2, n_dn = n - n_up;
  ftlm_op_t* op = make_sector(m, n_up, n_dn);
  const int dim = static_cast<int>(op_dim(op));
  const int lanczos_steps = std::min(lanczos_steps_req, dim);
  if (lanczos_steps_req > dim)
    std::cerr << "ftlm_bench_2x2: lanczos-steps=" << lanczos_steps_req << " capped to Hilbert dim=" << dim
              << " (no reorthogonalization)\n";
  std::cout << "=== ftlm_extended_hubbard: 2×2 (4-site) PBC benchmark ===\n";
  std::cout << "t=" << m.p.t << " t'=" << m.p.tp << " U=" << m.p.U << " V=" << m.p.V << "  sector N_up=" << n_up
            << " N_dn=" << n_dn << "  dim=" << dim << "\n";
  std::vector<double> H;
  auto t0 = clock_type::now();
  if (!dense_real(op, dim, &H)) {
    std::cerr << "ftlm_bench_2x2: expected real-symmetric H (set phi_x=phi_y=0)\n";
    return 2;
  }
  const double t_dense = seconds_since(t0);
  double e0 = 0.0, e1 = 0.0;
  ok(ftlm_hip_dense_jacobi(H.data(), dim, 1600, 1e-15, nullptr, &e0, &e1), "dense_jacobi");
  std::cout << "[reference] dense build " << (t_dense * 1e3) << " ms  jacobi_sweeps=1600  E_min=" << e0
            << "  E_max=" << e1 << "\n";
  std::vector<double> dense_full;
  if (plot_dir) {
    dense_full = sorted_jacobi(H, dim, 8000, 1e-12);
    std::cout << "[plot-data] dense spectrum: " << dense_full.size() << " eigenvalues (full Jacobi)\n";
  }
  {  // consistency: H_dense v against the device apply
    std::mt19937 rng2(99);
    std::normal_distribution<double> g2(0.0, 1.0);
    std::vector<cplx> x(static_cast<size_t>(dim)), ha(static_cast<size_t>(dim));
    for (int i = 0; i < dim; ++i) {
      const double re = g2(rng2);
      const double im = g2(rng2);
      x[static_cast<size_t>(i)] = cplx(re, im);
    }
    ok(ftlm_hip_apply_host(op, reinterpret_cast<const double*>(x.data()), reinterpret_cast<double*>(ha.data())),
       "apply_host");
    double err = 0.0;
    for (int i = 0; i < dim; ++i) {
      cplx s(0.0, 0.0);
      for (int j = 0; j < dim; ++j) s += H[static_cast<size_t>(i) * dim + j] * x[static_cast<size_t>(j)];
      err = std::max(err, std::abs(s - ha[static_cast<size_t>(i)]));
    }
    std::cout << "[check] max|H_dense v - apply_H v| = " << err << "\n";
    if (err > 1e-10) {
      std::cerr << "ftlm_bench_2x2: dense vs apply mismatch\n";
      return 3;
    }
  }
  // apply throughput: the reference's chain v <- H v on one vector, device resident
  {
    std::mt19937 rng(123);
    std::normal_distribution<double> gauss(0.0, 1.0);
    std::vector<cplx> v(static_cast<size_t>(dim));
    double nv = 0.0;
    for (int i = 0; i < dim; ++i) {
      const double re = gauss(rng);
      const double im = gauss(rng);
      v[static_cast<size_t>(i)] = cplx(re, im);
      nv += std::norm(v[static_cast<size_t>(i)]);
    }
    nv = std::sqrt(nv);
    for (auto& z : v) z /= nv;
    // the chain stays on the device: v, w are device buffers and every apply is ftlm_hip_apply
    double2* dv = nullptr;
    double2* dw = nullptr;
    if (hipMalloc(&dv, sizeof(double2) * dim) != hipSuccess || hipMalloc(&dw, sizeof(double2) * dim) != hipSuccess) {
      std::cerr << "ftlm_bench_2x2: hipMalloc failed\n";
      return 3;
    }
    if (hipMemcpy(dv, v.data(), sizeof(double2) * dim, hipMemcpyHostToDevice) != hipSuccess) fail("hipMemcpy");
    const int warmup = std::max(200, dim * 4);
    for (int i = 0; i < warmup; ++i) {
      ok(ftlm_hip_apply(op, dv, dw, 1, FTLM_C128, nullptr), "apply");
      std::swap(dv, dw);
    }
    if (hipDeviceSynchronize() != hipSuccess) fail("hipDeviceSynchronize");
    if (matvec_iters > 0) {
      t0 = clock_type::now();
      for (long long k = 0; k < matvec_iters; ++k) {
        ok(ftlm_hip_apply(op, dv, dw, 1, FTLM_C128, nullptr), "apply");
        std::swap(dv, dw);
      }
      if (hipDeviceSynchronize() != hipSuccess) fail("hipDeviceSynchronize");
      const double t_mat = seconds_since(t0);
      std::cout << "[bench] apply_H  iters=" << matvec_iters << "  time=" << (t_mat * 1e3)
                << " ms  throughput=" << (static_cast<double>(matvec_iters) / t_mat * 1e-6) << " Mapply/s\n";
    } else {
      std::cout << "[bench] apply_H  skipped (matvec-iters=0)\n";
    }
    (void)hipFree(dv);
    (void)hipFree(dw);
  }
  std::vector<double> alpha(static_cast<size_t>(lanczos_steps)), beta(static_cast<size_t>(lanczos_steps));
  int n_beta = 0, used = 0;
  t0 = clock_type::now();
  ok(ftlm_hip_lanczos(op, lanczos_steps, lanczos_seed, nullptr, alpha.data(), beta.data(), &n_beta, &used),
     "lanczos");
  const double t_lz = seconds_since(t0);
  double lo = 0.0, hi = 0.0;
  ok(ftlm_hip_tridiagonal_extrema(alpha.data(), used, beta.data(), used >= 2 ? used - 1 : 0, &lo, &hi),
     "tridiagonal_extrema");
  std::cout << "[bench] lanczos  steps=" << lanczos_steps << "  used=" << used << "  time=" << (t_lz * 1e3)
            << " ms  E_min≈" << lo << "  E_max≈" << hi << "  Δ_min=E_min(lz)-E_min(dense)=" << (lo - e0)
            << "  (no reorthogonalization)\n";
  if (plot_dir) {
    std::vector<double> T(static_cast<size_t>(used) * used, 0.0);
    for (int i = 0; i < used; ++i) T[static_cast<size_t>(i) * used + i] = alpha[static_cast<size_t>(i)];
    for (int i = 0; i < used - 1 && i < n_beta; ++i) {
      T[static_cast<size_t>(i) * used + i + 1] = beta[static_cast<size_t>(i)];
      T[static_cast<size_t>(i + 1) * used + i] = beta[static_cast<size_t>(i)];
    }
    const std::vector<double> ritz = sorted_jacobi(T, used, 6000, 1e-12);
    const std::string dir(plot_dir);
    std::filesystem::create_directories(dir);
    {
      std::ofstream f(dir + "/dense_evals.txt");
      for (double e : dense_full) f << e << "\n";
    }
    {
      std::ofstream f(dir + "/lanczos_ritz.txt");
      for (double e : ritz) f << e << "\n";
    }
    {
      std::ofstream f(dir + "/meta.txt");
      f << "Lx=" << m.p.Lx << " Ly=" << m.p.Ly << " t=" << m.p.t << " tp=" << m.p.tp << " U=" << m.p.U
        << " V=" << m.p.V << " phi_x=" << m.p.phi_x << " phi_y=" << m.p.phi_y << "\n";
      f << "N_up=" << n_up << " N_dn=" << n_dn << " dim_dense=" << dense_full.size()
        << " lanczos_krylov_dim=" << used << " n_ritz=" << ritz.size() << "\n";
    }
    std::cout << "[plot-data] wrote " << dir << "/{dense_evals.txt,lanczos_ritz.txt,meta.txt}\n";
  }
  ftlm_hip_op_destroy(op);
  return 0;
}
#elif FTLM_BENCH == 2
int run(int argc, char** argv) {
  const int samples = arg_int(argc, argv, "samples", 32);
  const int lanczos_steps = arg_int(argc, argv, "lanczos-steps", 128);
  const unsigned seed = static_cast<unsigned>(arg_int(argc, argv, "seed", 42));
  const int n_up = arg_int(argc, argv, "n-up", 4), n_dn = arg_int(argc, argv, "n-dn", 4);
  const Model m = make_model(4, 2, arg_double(argc, argv, "t", 1.0), arg_double(argc, argv, "tp", -0.35),
                             arg_double(argc, argv, "U", 5.75), arg_double(argc, argv, "V", 0.9));
  ftlm_op_t* op = make_sector(m, n_up, n_dn);
  const int64_t dim = op_dim(op);
  std::cout << "=== 4×2 FTLM precursor: stochastic Lanczos (matrix-free) ===\n";
  std::cout << "sector N_up=" << n_up << " N_dn=" << n_dn << " dim=" << dim << "  samples=" << samples
            << "  lanczos-steps=" << lanczos_steps << "\n";
  std::cout << "t=" << m.p.t << " t'=" << m.p.tp << " U=" << m.p.U << " V=" << m.p.V << "\n";
  const double rss0 = max_rss_mib();
  const auto t0 = clock_type::now();
  long long sum_steps = 0;
  double sum_lo = 0.0, sum_hi = 0.0;
  for (int s = 0; s < samples; ++s) {
    double lo = 0.0, hi = 0.0;
    int used = 0;
    ok(ftlm_hip_lanczos_extrema(op, lanczos_steps, seed + static_cast<unsigned>(s) * 7919u, &lo, &hi, &used),
       "lanczos_extrema");
    sum_steps += used;
    sum_lo += lo;
    sum_hi += hi;
  }
  const double wall_s = seconds_since(t0);
  std::cout << "avg Ritz E_min=" << (sum_lo / samples) << "  E_max=" << (sum_hi / samples)
            << " (not thermodynamic observables; diagnostic only)\n";
  // one H*v per Lanczos step (the reference counts calls of its apply)
  std::cout << "METRIC kind=FTLM_precursor_lanczos Lx=4 Ly=2 dim=" << dim << " n_up=" << n_up << " n_dn=" << n_dn
            << " samples=" << samples << " lanczos_steps_req=" << lanczos_steps << " sum_lanczos_steps_used="
            << sum_steps << " matvec_count=" << sum_steps << " wall_s=" << wall_s << " max_rss_mib=" << max_rss_mib()
            << " rss_baseline_mib=" << rss0 << "\n";
  ftlm_hip_op_destroy(op);
  return 0;
}
#else
int run(int argc, char** argv) {
  const double beta = arg_double(argc, argv, "beta", 20.0);
  const double mu_min = arg_double(argc, argv, "mu-min", -5.0);
  const double mu_max = arg_double(argc, argv, "mu-max", 25.0);
  const int n_mu = std::max(2, arg_int(argc, argv, "n-mu", 221));
  const char* out_path = find_arg(argc, argv, "out");
  const Model m = make_model(2, 2, arg_double(argc, argv, "t", 1.0), arg_double(argc, argv, "tp", -0.35),
                             arg_double(argc, argv, "U", 5.75), arg_double(argc, argv, "V", 0.9));
  const int n_sites = 4;
  std::cout << "=== 2×2 grand canonical: n(μ) at β=" << beta << " ===\n";
  std::cout << "t=" << m.p.t << " t'=" << m.p.tp << " U=" << m.p.U << " V=" << m.p.V << "  N_sites=" << n_sites
            << "\n";
  struct Level {
    double E;
    int N;
  };
  std::vector<Level> levels;
  const auto t0 = clock_type::now();
  for (int nu = 0; nu <= n_sites; ++nu)
    for (int nd = 0; nd <= n_sites; ++nd) {
      ftlm_op_t* op = make_sector(m, nu, nd);
      const int dim = static_cast<int>(op_dim(op));
      if (dim > 0) {
        std::vector<double> H;
        if (!dense_real(op, dim, &H)) {
          std::cerr << "ftlm_bench_2x2_nmu: expected real-symmetric H\n";
          return 2;
        }
        for (double e : sorted_jacobi(H, dim, 12000, 1e-14)) levels.push_back(Level{e, nu + nd});
      }
      ftlm_hip_op_destroy(op);
    }
  const double t_ed = seconds_since(t0);
  std::cout << "exact diagonalization: " << levels.size() << " eigenlevels in " << (t_ed * 1e3) << " ms\n";
  std::ostream* out = &std::cout;
  std::ofstream file;
  if (out_path) {
    file.open(out_path);
    if (!file) {
      std::cerr << "ftlm_bench_2x2_nmu: cannot open " << out_path << "\n";
      return 2;
    }
    out = &file;
  }
  *out << "# 2x2 PBC grand canonical  mu  n=<N>/N_sites  beta=" << beta << "  t=" << m.p.t << " tp=" << m.p.tp
       << " U=" << m.p.U << " V=" << m.p.V << "\n";
  for (int i = 0; i < n_mu; ++i) {  // grand_canonical_expectation (bench_2x2_nmu.cpp:133-156)
    const double mu = mu_min + (static_cast<double>(i) / static_cast<double>(n_mu - 1)) * (mu_max - mu_min);
    double mx = -std::numeric_limits<double>::infinity();
    for (const Level& lv : levels) mx = std::max(mx, -beta * (lv.E - mu * static_cast<double>(lv.N)));
    double sw = 0.0, snw = 0.0;
    for (const Level& lv : levels) {
      const double w = std::exp(-beta * (lv.E - mu * static_cast<double>(lv.N)) - mx);
      sw += w;
      snw += static_cast<double>(lv.N) * w;
    }
    *out << mu << "\t" << (snw / sw) / static_cast<double>(n_sites) << "\n";
  }
  if (out_path) std::cout << "wrote " << out_path << "\n";
  return 0;
}
#endif

}  // namespace

int main(int argc, char** argv) { return run(argc, argv); }
