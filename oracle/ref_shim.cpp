// ref_shim.cpp -- C entry points over the reference's own hot-path sources, compiled in this
// container from /root/reference (see oracle/Makefile) into oracle/_ref/libftlm_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py and make_nmu_golden.py to write
// the golden fixtures against which tests/test_oracle.py pins the oracle restatement (and the GPU
// tests the device path).  It calls the reference API exactly
// as the reference's drivers do; it contains no numerics of its own except argument marshalling.

#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ftlm/fock_basis.hpp"
#include "ftlm/ftlm_thermo.hpp"
#include "ftlm/hubbard_hamiltonian.hpp"
#include "ftlm/hubbard_params.hpp"
#include "ftlm/lanczos.hpp"
#include "ftlm/lattice.hpp"
#include "ftlm/symmetry/orbit_k_block_basis.hpp"
#include "ftlm/version.hpp"

namespace {

ftlm::HubbardParams params_from(const double* p) {
  ftlm::HubbardParams h;
  h.Lx = static_cast<int>(p[0]);
  h.Ly = static_cast<int>(p[1]);
  h.t = p[2];
  h.tp = p[3];
  h.U = p[4];
  h.V = p[5];
  h.phi_x = p[6];
  h.phi_y = p[7];
  return h;
}

struct Geo {
  std::vector<ftlm::SpinfulHopping> hops;
  std::vector<ftlm::NearestPair> pairs;
};

Geo geo_of(const ftlm::HubbardParams& p) {
  Geo g;
  ftlm::RectLattice lat{p.Lx, p.Ly};
  ftlm::build_hubbard_geometry(p, lat, &g.hops, &g.pairs);
  return g;
}

}  // namespace

extern "C" {

const char* ref_version() { return ftlm::version_string(); }

int ref_geometry(const double* p8, int* from, int* to, int* spin, double* re, double* im, int cap,
                 int* nh, int* pi, int* pj, int pcap, int* np) {
  const Geo g = geo_of(params_from(p8));
  *nh = static_cast<int>(g.hops.size());
  *np = static_cast<int>(g.pairs.size());
  if (*nh > cap || *np > pcap) return 1;
  for (int k = 0; k < *nh; ++k) {
    from[k] = g.hops[k].from;
    to[k] = g.hops[k].to;
    spin[k] = g.hops[k].spin;
    re[k] = g.hops[k].coeff.real();
    im[k] = g.hops[k].coeff.imag();
  }
  for (int k = 0; k < *np; ++k) {
    pi[k] = g.pairs[k].i;
    pj[k] = g.pairs[k].j;
  }
  return 0;
}

int ref_basis_masks(int n, int nu, int nd, uint64_t* up, uint64_t* dn, int cap) {
  try {
    ftlm::FockBasis b(n, nu, nd);
    if (b.dim() > cap) return -2;
    for (int i = 0; i < b.dim(); ++i) {
      up[i] = b.up_mask(i);
      dn[i] = b.down_mask(i);
    }
    return b.dim();
  } catch (...) {
    return -1;
  }
}

int ref_index_of(int n, int nu, int nd, const uint64_t* up, const uint64_t* dn, int cnt, int* out) {
  try {
    ftlm::FockBasis b(n, nu, nd);
    for (int i = 0; i < cnt; ++i) out[i] = b.index_of(up[i], dn[i]);
    return 0;
  } catch (...) {
    return -1;
  }
}

int ref_apply(const double* p8, int nu, int nd, const double* x, double* y) {
  const ftlm::HubbardParams p = params_from(p8);
  const Geo g = geo_of(p);
  ftlm::FockBasis b(p.Lx * p.Ly, nu, nd);
  ftlm::apply_extended_hubbard(p, b, g.hops, g.pairs, reinterpret_cast<const std::complex<double>*>(x),
                               reinterpret_cast<std::complex<double>*>(y));
  return 0;
}

int ref_lanczos(const double* p8, int nu, int nd, int M, unsigned seed, double* alpha, double* beta, int* nb) {
  const ftlm::HubbardParams p = params_from(p8);
  const Geo g = geo_of(p);
  ftlm::FockBasis b(p.Lx * p.Ly, nu, nd);
  auto ap = [&](const std::complex<double>* x, std::complex<double>* y) {
    ftlm::apply_extended_hubbard(p, b, g.hops, g.pairs, x, y);
  };
  std::vector<double> a, be;
  const int used = ftlm::lanczos_tridiagonal(b.dim(), ap, M, seed, &a, &be);
  std::memcpy(alpha, a.data(), sizeof(double) * a.size());
  std::memcpy(beta, be.data(), sizeof(double) * be.size());
  *nb = static_cast<int>(be.size());
  return used;
}

int ref_lanczos_extrema(const double* p8, int nu, int nd, int M, unsigned seed, double* lo, double* hi) {
  const ftlm::HubbardParams p = params_from(p8);
  const Geo g = geo_of(p);
  ftlm::FockBasis b(p.Lx * p.Ly, nu, nd);
  auto ap = [&](const std::complex<double>* x, std::complex<double>* y) {
    ftlm::apply_extended_hubbard(p, b, g.hops, g.pairs, x, y);
  };
  const ftlm::LanczosExtrema e = ftlm::lanczos_extrema(b.dim(), ap, M, seed);
  *lo = e.min_eval;
  *hi = e.max_eval;
  return e.steps_used;
}

int ref_tridiagonal_extrema(const double* alpha, int n, const double* beta, int nb, double* lo, double* hi) {
  std::vector<double> a(alpha, alpha + n), b(beta, beta + nb);
  const auto r = ftlm::tridiagonal_extrema(a, b, n);
  *lo = r.first;
  *hi = r.second;
  return 0;
}

// One-beta FTLM lnZ.  real=1: ftlm_log_partition_real with the apply_real wrapper used by
// benchmarks/bench_ftlm_nmu_rect.cpp:151-159.
double ref_ftlm_log_partition(const double* p8, int nu, int nd, double bt, int R, int M, unsigned seed, int real) {
  const ftlm::HubbardParams p = params_from(p8);
  const Geo g = geo_of(p);
  ftlm::FockBasis b(p.Lx * p.Ly, nu, nd);
  const int dim = b.dim();
  ftlm::FtlmParams fp;
  fp.n_random = R;
  fp.lanczos_steps = M;
  fp.seed = seed;
  if (real) {
    std::vector<std::complex<double>> xc(static_cast<size_t>(dim)), yc(static_cast<size_t>(dim));
    auto apply_real = [&](const double* x, double* y) {
      for (int i = 0; i < dim; ++i) xc[i] = {x[i], 0.0};
      ftlm::apply_extended_hubbard(p, b, g.hops, g.pairs, xc.data(), yc.data());
      for (int i = 0; i < dim; ++i) y[i] = yc[i].real();
    };
    return ftlm::ftlm_log_partition_real(dim, apply_real, bt, fp);
  }
  auto ap = [&](const std::complex<double>* x, std::complex<double>* y) {
    ftlm::apply_extended_hubbard(p, b, g.hops, g.pairs, x, y);
  };
  return ftlm::ftlm_log_partition_complex(dim, ap, bt, fp);
}

// Generic operator given as a dense complex matrix (row-major): ftlm_log_partition_complex on it.
double ref_ftlm_log_partition_dense(const double* H, int dim, double bt, int R, int M, unsigned seed) {
  ftlm::FtlmParams fp;
  fp.n_random = R;
  fp.lanczos_steps = M;
  fp.seed = seed;
  auto ap = [&](const std::complex<double>* x, std::complex<double>* y) {
    for (int i = 0; i < dim; ++i) {
      std::complex<double> s(0.0, 0.0);
      for (int j = 0; j < dim; ++j)
        s += std::complex<double>(H[2 * (i * dim + j)], H[2 * (i * dim + j) + 1]) * x[j];
      y[i] = s;
    }
  };
  return ftlm::ftlm_log_partition_complex(dim, ap, bt, fp);
}

int ref_orbit_basis(int lx, int ly, int nu, int nd, int kx, int ky, uint32_t* reps, int* osz, int* ssz,
                    double* A, int cap) {
  ftlm::FockBasis fb(lx * ly, nu, nd);
  ftlm::symmetry::OrbitKBlockBasis ob;
  ob.build(fb, ftlm::symmetry::MomentumSector{kx, ky, lx, ly});
  const int dk = ob.dim();
  if (dk > cap) return -2;
  for (int r = 0; r < dk; ++r) {
    const auto& e = ob.entries()[static_cast<size_t>(r)];
    reps[r] = (static_cast<uint32_t>(e.rep.up) << 16) | e.rep.dn;
    osz[r] = e.orbit_size;
    ssz[r] = e.stabilizer_size;
    A[2 * r] = e.A_k.real();
    A[2 * r + 1] = e.A_k.imag();
  }
  return dk;
}

int ref_orbit_map(int lx, int ly, int nu, int nd, int kx, int ky, const uint32_t* st, int cnt, int* row,
                  int* dx, int* dy, double* eta) {
  ftlm::FockBasis fb(lx * ly, nu, nd);
  ftlm::symmetry::OrbitKBlockBasis ob;
  ob.build(fb, ftlm::symmetry::MomentumSector{kx, ky, lx, ly});
  for (int i = 0; i < cnt; ++i) {
    ftlm::symmetry::RawState s{static_cast<uint16_t>(st[i] >> 16), static_cast<uint16_t>(st[i] & 0xFFFFu)};
    int r = -1, a = 0, c = 0;
    double e = 0.0;
    if (!ob.map_raw_to_rep(s, &r, &a, &c, &e)) r = -1;
    row[i] = r;
    dx[i] = a;
    dy[i] = c;
    eta[i] = e;
  }
  return 0;
}

int ref_orbit_apply(const double* p8, int nu, int nd, int kx, int ky, const double* x, double* y) {
  const ftlm::HubbardParams p = params_from(p8);
  const Geo g = geo_of(p);
  ftlm::FockBasis fb(p.Lx * p.Ly, nu, nd);
  ftlm::symmetry::OrbitKBlockBasis ob;
  ob.build(fb, ftlm::symmetry::MomentumSector{kx, ky, p.Lx, p.Ly});
  ftlm::symmetry::apply_orbit_k_block(p, fb, ob, g.hops, g.pairs,
                                      reinterpret_cast<const std::complex<double>*>(x),
                                      reinterpret_cast<std::complex<double>*>(y));
  return ob.dim();
}

double ref_orbit_ftlm_log_partition(const double* p8, int nu, int nd, int kx, int ky, double bt, int R, int M,
                                    unsigned seed) {
  const ftlm::HubbardParams p = params_from(p8);
  const Geo g = geo_of(p);
  ftlm::FockBasis fb(p.Lx * p.Ly, nu, nd);
  ftlm::symmetry::OrbitKBlockBasis ob;
  ob.build(fb, ftlm::symmetry::MomentumSector{kx, ky, p.Lx, p.Ly});
  auto ap = [&](const std::complex<double>* x, std::complex<double>* y) {
    ftlm::symmetry::apply_orbit_k_block(p, fb, ob, g.hops, g.pairs, x, y);
  };
  ftlm::FtlmParams fp;
  fp.n_random = R;
  fp.lanczos_steps = M;
  fp.seed = seed;
  return ftlm::ftlm_log_partition_complex(ob.dim(), ap, bt, fp);
}

}  // extern "C"
