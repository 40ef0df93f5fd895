// Host-only checks run under AddressSanitizer + UndefinedBehaviorSanitizer and under
// ThreadSanitizer (tests/sanitize/Makefile).  They cover the host code that needs no GPU:
//   - the exact parallel draw of the reference Gaussian starts (ftlm_extended_hubbard_amd/csrc/
//     gauss_mt.hpp): MT19937 against std::mt19937, jump-ahead against discard(), and the
//     multi-threaded draw against the oracle's serial mt19937 / normal_distribution stream
//     (src/lanczos.cpp:140-168), bit for bit;
//   - the drivers' host pieces (csrc/host/drivers/driver_host.hpp): --key=value parsing and the
//     n(mu) mixing against the oracle's gc_density (bench_ftlm_nmu_rect.cpp:186-219);
//   - the CPU oracle itself (oracle/ftlm_oracle.cpp, compiled into this program): apply,
//     Lanczos, FTLM lnZ, orbit blocks and the dense helpers on small sectors, with identities that
//     a memory or arithmetic fault would break (Hermiticity, real spectra, block dims).
// Exit code 0 and the line "host_checks ok" on success.

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "drivers/driver_host.hpp"
#include "gauss_mt.hpp"

extern "C" {
long orc_basis_dim(int n, int nu, int nd);
int orc_apply(const double* params8, int nu, int nd, const double* x, double* y);
int orc_gauss_start(long dim, unsigned seed, double* q);
int orc_lanczos(const double* params8, int nu, int nd, int M, unsigned seed, const double* q0, double* alpha,
                double* beta, int* nb);
double orc_ftlm_log_partition(const double* params8, int nu, int nd, double bt, int R, int M, unsigned seed, int real);
void orc_gc_density(const double* lnZ, const int* N, int nsec, double bt, const double* mu, int nmu, int n_sites,
                    double* n_out);
int orc_orbit_basis(int lx, int ly, int nu, int nd, int kx, int ky, double tol, uint32_t* reps, int* orbit_size,
                    int* stab_size, double* A, long cap);
int orc_orbit_apply(const double* params8, int nu, int nd, int kx, int ky, double tol, const double* x, double* y);
int orc_dense_hamiltonian(const double* params8, int nu, int nd, double* H);
double orc_exact_log_partition(const double* H, int d, double bt);
}

namespace {

int failures = 0;
#define EXPECT(cond)                                                      \
  do {                                                                    \
    if (!(cond)) {                                                        \
      std::fprintf(stderr, "%s:%d: failed: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                         \
    }                                                                     \
  } while (0)

using cplx = std::complex<double>;

// The threaded draw against the oracle's serial stream: T threads, several rounds of seg words.
void check_gauss_threads(long dim, int T, int seg) {
  std::vector<double> want(2 * static_cast<size_t>(dim));
  orc_gauss_start(dim, 4242u, want.data());
  std::vector<cplx> got(static_cast<size_t>(dim));
  long calls = 0, next = 0;
  bool ordered = true;
  ftlm_gauss::draw_sample(
      4242u, dim, T,
      [&](int64_t first, int64_t n, const cplx* v) {
        ordered = ordered && first == next;
        next = first + n;
        ++calls;
        std::memcpy(got.data() + first, v, static_cast<size_t>(n) * sizeof(cplx));
      },
      seg);
  EXPECT(ordered && next == dim && calls > T);  // more than one round
  EXPECT(std::memcmp(got.data(), want.data(), want.size() * sizeof(double)) == 0);
}

void check_gauss() {
  using namespace ftlm_gauss;
  {
    Mt19937 a(123);
    std::mt19937 b(123);
    std::vector<uint32_t> w(3000);
    a.fill(w.data(), 7);
    for (int i = 0; i < 7; ++i) EXPECT(w[i] == b());
    a.fill(w.data(), 3000);  // partial block, then whole blocks, then a tail
    for (int i = 0; i < 3000; ++i) EXPECT(w[i] == b());
  }
  const Poly phi = char_poly();
  EXPECT(poly_deg(phi) == kDeg);
  const int dphi = poly_deg(phi);
  for (uint64_t J : {uint64_t{4099}, uint64_t{4} * kSegWords + 4}) {
    Mt19937 a(77);
    a.apply_poly(pow_t(J, phi, dphi));
    std::mt19937 b(77);
    b.discard(J);
    bool same = true;
    for (int i = 0; i < 2000; ++i) same = same && a() == b();
    EXPECT(same);
  }
  check_gauss_threads(3500000, 3, kSegWords);  // two rounds of the default segment
}

void check_driver_host() {
  using namespace ftlm_driver;
  const char* argv_c[] = {"prog", "--n-mu=41", "--beta=2.5", "--n-mu-extra=3", "--out=", "--seed=-3", "--n-mu=7", nullptr};
  char** argv = const_cast<char**>(argv_c);
  const int argc = 7;
  EXPECT(arg_int(argc, argv, "n-mu", 0) == 41);  // first match
  EXPECT(arg_double(argc, argv, "beta", 0.0) == 2.5);
  EXPECT(arg_int(argc, argv, "missing", 17) == 17);
  EXPECT(has_arg(argc, argv, "out") && std::strlen(find_arg(argc, argv, "out")) == 0);
  EXPECT(arg_int(argc, argv, "seed", 0) == -3);
  EXPECT(!has_arg(argc, argv, "n"));
  EXPECT(logsumexp2(-INFINITY, 1.5) == 1.5 && logsumexp2(2.0, -INFINITY) == 2.0);
  EXPECT(std::abs(logsumexp2(0.0, 0.0) - std::log(2.0)) < 1e-15);
  // n(mu) mixing against the oracle's gc_density, with -inf sectors; and an all -inf case
  std::vector<double> lz = {1.0, -INFINITY, 3.5, 2.25, -INFINITY, 0.5};
  std::vector<int> N = {0, 1, 2, 3, 4, 5};
  const int nmu = 9;
  std::ostringstream os;
  write_nmu(os, lz, N, 4, 3.0, -2.0, 6.0, nmu);
  std::vector<double> mu(nmu), want(nmu);
  for (int k = 0; k < nmu; ++k) mu[k] = -2.0 + (static_cast<double>(k) / (nmu - 1)) * 8.0;
  orc_gc_density(lz.data(), N.data(), 6, 3.0, mu.data(), nmu, 4, want.data());
  std::istringstream is(os.str());
  for (int k = 0; k < nmu; ++k) {
    double m = 0.0, n = 0.0;
    is >> m >> n;
    EXPECT(std::abs(n - want[k]) <= 1e-5 * std::max(1.0, std::abs(want[k])));
  }
  std::vector<double> none(3, -INFINITY);
  std::vector<int> N3 = {1, 2, 3};
  std::ostringstream os2;
  write_nmu(os2, none, N3, 2, 1.0, 0.0, 1.0, 2);
  EXPECT(os2.str() == "0\t0\n1\t0\n");
}

void check_oracle() {
  const double p8[8] = {3, 2, 1.0, -0.3, 5.0, 0.7, 0.2, -0.1};
  const int nu = 3, nd = 2;
  const long d = orc_basis_dim(6, nu, nd);
  EXPECT(d == 20 * 15);
  std::mt19937 rng(3);
  std::normal_distribution<double> g(0.0, 1.0);
  std::vector<cplx> x(static_cast<size_t>(d)), z(static_cast<size_t>(d)), hx(static_cast<size_t>(d)),
      hz(static_cast<size_t>(d));
  for (long i = 0; i < d; ++i) {
    x[i] = cplx(g(rng), g(rng));
    z[i] = cplx(g(rng), g(rng));
  }
  EXPECT(orc_apply(p8, nu, nd, reinterpret_cast<double*>(x.data()), reinterpret_cast<double*>(hx.data())) == 0);
  EXPECT(orc_apply(p8, nu, nd, reinterpret_cast<double*>(z.data()), reinterpret_cast<double*>(hz.data())) == 0);
  cplx a(0, 0), b(0, 0);
  for (long i = 0; i < d; ++i) {
    a += std::conj(x[i]) * hz[i];
    b += std::conj(hx[i]) * z[i];
  }
  EXPECT(std::abs(a - b) <= 1e-10 * std::abs(a));  // Hermitian
  std::vector<double> al(40), be(40);
  int nb = 0;
  const int used = orc_lanczos(p8, nu, nd, 40, 7u, nullptr, al.data(), be.data(), &nb);
  EXPECT(used == 40 && nb >= 39);
  for (int k = 0; k < nb; ++k) EXPECT(be[k] >= 0.0);
  // FTLM lnZ at beta -> 0 approaches ln dim
  const double lz0 = orc_ftlm_log_partition(p8, nu, nd, 1e-9, 4, 20, 11u, 0);
  EXPECT(std::abs(lz0 - std::log(static_cast<double>(d))) < 1e-6);
  // dense H: exact lnZ at small beta against the trace expansion ln d - beta tr H / d
  std::vector<double> H(2 * static_cast<size_t>(d) * d);
  EXPECT(orc_dense_hamiltonian(p8, nu, nd, H.data()) == 0);
  double tr = 0.0;
  for (long i = 0; i < d; ++i) tr += H[2 * (static_cast<size_t>(i) * d + i)];
  const double bt = 1e-6;
  const double ex = orc_exact_log_partition(H.data(), static_cast<int>(d), bt);
  EXPECT(std::abs(ex - (std::log(static_cast<double>(d)) - bt * tr / d)) < 1e-9);
  // orbit blocks of 3x2 (3,3) at zero flux: block dims sum to the sector dimension
  const double q8[8] = {3, 2, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0};
  long total = 0;
  for (int ky = 0; ky < 2; ++ky)
    for (int kx = 0; kx < 3; ++kx) {
      std::vector<uint32_t> reps(400);
      std::vector<int> os(400), ss(400);
      std::vector<double> A(800);
      const int n = orc_orbit_basis(3, 2, 3, 3, kx, ky, 1e-12, reps.data(), os.data(), ss.data(), A.data(), 400);
      EXPECT(n >= 0);
      total += n;
      if (n > 0) {
        std::vector<cplx> v(static_cast<size_t>(n), cplx(1.0, 0.5)), w(static_cast<size_t>(n));
        EXPECT(orc_orbit_apply(q8, 3, 3, kx, ky, 1e-12, reinterpret_cast<double*>(v.data()),
                               reinterpret_cast<double*>(w.data())) == 0);
      }
    }
  EXPECT(total == 400);
}

}  // namespace

// With the argument "threads" (the ThreadSanitizer run) only the threaded pieces run, on short
// segments so that many rounds fit in a few seconds of instrumented time.
int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "threads") == 0) {
    check_gauss_threads(150000, 4, 1 << 14);
    check_gauss_threads(40000, 3, 1 << 12);
    check_driver_host();
  } else {
    check_gauss();
    check_driver_host();
    check_oracle();
  }
  if (failures) {
    std::fprintf(stderr, "host_checks: %d failure(s)\n", failures);
    return 1;
  }
  std::printf("host_checks ok\n");
  return 0;
}
