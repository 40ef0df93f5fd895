// grad.hip -- exact derivatives of the FTLM lnZ estimator with respect to Hamiltonian parameters
// (forward-mode / tangent Lanczos).
//
// The estimator is the one of ftlm_log_partition_complex (src/ftlm_thermo.cpp:213-256) with the
// converged Gauss rule: lnZ = ln dim + log-mean-exp_r L_r, L_r = ln(e_1^T exp(-beta T_r) e_1),
// T_r the Lanczos tridiagonal of sample r.  For a parameter theta with H(theta) and dH = dH/dtheta
// (an operator of the same kind: derivative hop coefficients, or U/V diagonal terms), each step
// of the reference recurrence is differentiated exactly:
//     w = H q                              dw  = dH q + H dq
//     z = <q|w>, alpha = Re z              dz  = <dq|w> + <q|dw>,  dalpha = Re dz
//     w -= z q + beta' q'                  dw -= dz q + z dq + dbeta' q' + beta' dq'
//     beta = |w|                           dbeta = Re<w|dw> / beta
//     q_new = w / beta                     dq_new = (dw - dbeta q_new) / beta
// (the start vector does not depend on theta).  The derivative of L_r follows from the
// eigen-decomposition T = U diag(lambda) U^T:
//     dL_r = (1/Z_r) sum_ij u_i u_j g[lambda_i, lambda_j] (U^T dT U)_ij,   u = U^T e_1,
// with g[.,.] the divided difference of exp(-beta lambda) (the Frechet derivative of exp(-beta T)),
// contracted with dT = tridiag(dalpha, dbeta); dlnZ = sum_r softmax(L_r) dL_r.  All FP64.

#include <algorithm>
#include <cmath>
#include <complex>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"

namespace ftlm_hip {

void gauss_mt(int64_t dim, unsigned seed, std::vector<std::complex<double>>* q);

namespace {

constexpr int kGBlock = 256;
constexpr int kGMaxGrid = 2048;

// Per-block partial sums of <a_r|b_r> over the flat [dim][R] layout (thread t keeps vector t % R).
__global__ __launch_bounds__(kGBlock) void k_gdot(int64_t n, int R, const double2* __restrict__ a,
                                                  const double2* __restrict__ b, double2* __restrict__ part) {
  const int nthr = blockDim.x;
  double sr = 0.0, si = 0.0;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * nthr + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * nthr) {
    const double2 x = a[f], y = b[f];
    sr += x.x * y.x + x.y * y.y;
    si += x.x * y.y - x.y * y.x;
  }
  __shared__ double rr[kGBlock], ri[kGBlock];
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  if (threadIdx.x < R) {
    double x = 0.0, y = 0.0;
    for (int t = threadIdx.x; t < nthr; t += R) {
      x += rr[t];
      y += ri[t];
    }
    part[static_cast<int64_t>(blockIdx.x) * R + threadIdx.x] = make_double2(x, y);
  }
}

__global__ void k_greduce(const double2* __restrict__ part, int nblk, int R, double2* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double x = 0.0, y = 0.0;
  for (int b = 0; b < nblk; ++b) {
    x += part[static_cast<int64_t>(b) * R + r].x;
    y += part[static_cast<int64_t>(b) * R + r].y;
  }
  out[r] = make_double2(x, y);
}

// dst += src
__global__ void k_gadd(int64_t n, double2* __restrict__ dst, const double2* __restrict__ src) {
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double2 a = dst[f], b = src[f];
    dst[f] = make_double2(a.x + b.x, a.y + b.y);
  }
}

// w -= z q + bp qp   (coefficients per vector r = f % R; dead vectors are zeroed)
__global__ void k_gupd_w(int64_t n, int R, double2* __restrict__ w, const double2* __restrict__ q,
                         const double2* __restrict__ qp, const double2* __restrict__ z, const double* __restrict__ bp,
                         const int* __restrict__ alive, int use_prev) {
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int r = static_cast<int>(f % R);
    double2 c = w[f];
    const double2 a = q[f], zz = z[r];
    c.x -= zz.x * a.x - zz.y * a.y;
    c.y -= zz.x * a.y + zz.y * a.x;
    if (use_prev) {
      const double2 b = qp[f];
      c.x -= bp[r] * b.x;
      c.y -= bp[r] * b.y;
    }
    if (!alive[r]) c = make_double2(0.0, 0.0);
    w[f] = c;
  }
}

// dw -= dz q + z dq + dbp qp + bp dqp
__global__ void k_gupd_dw(int64_t n, int R, double2* __restrict__ dw, const double2* __restrict__ q,
                          const double2* __restrict__ dq, const double2* __restrict__ qp, const double2* __restrict__ dqp,
                          const double2* __restrict__ z, const double2* __restrict__ dz, const double* __restrict__ bp,
                          const double* __restrict__ dbp, const int* __restrict__ alive, int use_prev) {
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int r = static_cast<int>(f % R);
    double2 c = dw[f];
    const double2 a = q[f], da = dq[f], zz = z[r], dzz = dz[r];
    c.x -= dzz.x * a.x - dzz.y * a.y + zz.x * da.x - zz.y * da.y;
    c.y -= dzz.x * a.y + dzz.y * a.x + zz.x * da.y + zz.y * da.x;
    if (use_prev) {
      const double2 b = qp[f], db = dqp[f];
      c.x -= dbp[r] * b.x + bp[r] * db.x;
      c.y -= dbp[r] * b.y + bp[r] * db.y;
    }
    if (!alive[r]) c = make_double2(0.0, 0.0);
    dw[f] = c;
  }
}

// q_new = w / beta (in place in w)
__global__ void k_gnorm_q(int64_t n, int R, double2* __restrict__ w, const double* __restrict__ beta,
                          const int* __restrict__ alive) {
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int r = static_cast<int>(f % R);
    const double2 c = w[f];
    w[f] = alive[r] ? make_double2(c.x / beta[r], c.y / beta[r]) : make_double2(0.0, 0.0);
  }
}

// dq_new = (dw - dbeta q_new) / beta (in place in dw)
__global__ void k_gnorm_dq(int64_t n, int R, double2* __restrict__ dw, const double2* __restrict__ qn,
                           const double* __restrict__ beta, const double* __restrict__ dbeta,
                           const int* __restrict__ alive) {
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int r = static_cast<int>(f % R);
    const double2 c = dw[f], a = qn[f];
    dw[f] = alive[r] ? make_double2((c.x - dbeta[r] * a.x) / beta[r], (c.y - dbeta[r] * a.y) / beta[r])
                     : make_double2(0.0, 0.0);
  }
}

__global__ void k_gphase(int64_t dim, int R, const uint32_t* ids, double2* q) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < dim;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    for (int r = 0; r < R; ++r) {
      uint64_t z = 0x5eedull * 0x9e3779b97f4a7c15ull + ids[r];
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      z = z ^ (z >> 31);
      z += static_cast<uint64_t>(e);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      z = z ^ (z >> 31);
      const double th = 6.283185307179586 * static_cast<double>(z >> 11) * 0x1.0p-53;
      double s, c;
      sincos(th, &s, &c);
      q[e * R + r] = make_double2(c, s);
    }
}

// Per sample s (one thread): converged eigen-decomposition of T_s (implicit QL, eigenvectors in
// `U`, n x n), then for every beta: L_s(beta) and dL_s/dtheta_p.
// alpha/beta/dalpha/dbeta: [s][Mmax] (dalpha/dbeta: [p][s][Mmax]).  Outputs L [s][nb], dL [p][s][nb].
__global__ void k_grad_tridiag(const double* __restrict__ alpha, const double* __restrict__ beta,
                               const double* __restrict__ dalpha, const double* __restrict__ dbeta,
                               const int* __restrict__ nvec, int nsamp, int Mmax, int npar,
                               const double* __restrict__ betas, int nb, double* __restrict__ Uw,
                               double* __restrict__ W, double* __restrict__ lam_w, double* __restrict__ L,
                               double* __restrict__ dL, int* __restrict__ status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsamp) return;
  const int n = nvec[s];
  if (n <= 0) return;
  double* U = Uw + static_cast<int64_t>(s) * Mmax * Mmax;   // U[a * n + i]: component a of eigenvector i
  double* Wm = W + static_cast<int64_t>(s) * Mmax * Mmax;   // u_i u_j g[lambda_i, lambda_j]
  double* d = lam_w + static_cast<int64_t>(s) * 2 * Mmax;
  double* e = d + Mmax;
  for (int i = 0; i < n; ++i) {
    d[i] = alpha[static_cast<int64_t>(s) * Mmax + i];
    e[i] = i < n - 1 ? beta[static_cast<int64_t>(s) * Mmax + i] : 0.0;
    for (int a = 0; a < n; ++a) U[a * n + i] = (a == i) ? 1.0 : 0.0;
  }
  int bad = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 1e-300 + 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) {
        bad = 1;
        break;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double sn = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool deflated = false;
      for (; i >= l; --i) {
        double f = sn * e[i];
        const double b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        sn = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * sn + 2.0 * c * b;
        p = sn * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int a = 0; a < n; ++a) {  // rotate columns i, i+1 of the eigenvector matrix
          const double ua = U[a * n + i], ub = U[a * n + i + 1];
          U[a * n + i + 1] = sn * ua + c * ub;
          U[a * n + i] = c * ua - sn * ub;
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  if (bad) atomicAdd(status, 1);
  for (int ib = 0; ib < nb; ++ib) {
    const double bt = betas[ib];
    // L = ln sum_i u_i^2 exp(-bt lambda_i), shifted by the largest exponent
    double mx = -INFINITY;
    for (int i = 0; i < n; ++i) {
      const double w = U[i] * U[i];
      if (w > 0.0) mx = fmax(mx, -bt * d[i] + log(w));
    }
    double z = 0.0;
    for (int i = 0; i < n; ++i) {
      const double w = U[i] * U[i];
      if (w > 0.0) z += exp(-bt * d[i] + log(w) - mx);
    }
    L[static_cast<int64_t>(s) * nb + ib] = mx + log(z);
    // W_ij = u_i u_j g[lambda_i, lambda_j] / Z with g(lambda) = exp(-bt lambda) (same shift)
    const double lz = mx + log(z);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        const double ui = U[i], uj = U[j];
        double v = 0.0;
        if (ui != 0.0 && uj != 0.0) {
          // divided difference of exp(-bt lambda) / Z: close eigenvalues through
          // -bt exp(-bt mean) sinh(x)/x, x = bt (lambda_i - lambda_j) / 2; otherwise the plain quotient
          const double xd = 0.5 * bt * (d[i] - d[j]);
          if (fabs(xd) < 1e-3) {
            const double shc = 1.0 + xd * xd / 6.0;
            v = ui * uj * (-bt) * exp(-bt * 0.5 * (d[i] + d[j]) - lz) * shc;
          } else {
            v = ui * uj * (exp(-bt * d[i] - lz) - exp(-bt * d[j] - lz)) / (d[i] - d[j]);
          }
        }
        Wm[i * n + j] = v;
        Wm[j * n + i] = v;
      }
    // G_kk and G_{k,k+1} (G = U W U^T), contracted with dT for every parameter
    for (int p = 0; p < npar; ++p) dL[(static_cast<int64_t>(p) * nsamp + s) * nb + ib] = 0.0;
    for (int k = 0; k < n; ++k) {
      double gkk = 0.0, gk1 = 0.0;
      for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int j = 0; j < n; ++j) t += Wm[i * n + j] * U[k * n + j];
        gkk += U[k * n + i] * t;
        if (k + 1 < n) gk1 += U[(k + 1) * n + i] * t;
      }
      for (int p = 0; p < npar; ++p) {
        const int64_t o = (static_cast<int64_t>(p) * nsamp + s) * Mmax + k;
        double acc = gkk * dalpha[o];
        if (k + 1 < n) acc += 2.0 * gk1 * dbeta[o];
        dL[(static_cast<int64_t>(p) * nsamp + s) * nb + ib] += acc;
      }
    }
  }
}

// lnZ(beta) = ln dim + log-mean-exp_s L_s; dlnZ_p = sum_s softmax(L)_s dL_s,p.
__global__ void k_grad_combine(const double* L, const double* dL, const int* nvec, int nsamp, int npar, int nb,
                               double lndim, double* lnZ, double* grad) {
  const int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nb) return;
  double mx = -INFINITY;
  int cnt = 0;
  for (int s = 0; s < nsamp; ++s)
    if (nvec[s] > 0) {
      mx = fmax(mx, L[static_cast<int64_t>(s) * nb + ib]);
      ++cnt;
    }
  double z = 0.0;
  for (int s = 0; s < nsamp; ++s)
    if (nvec[s] > 0) z += exp(L[static_cast<int64_t>(s) * nb + ib] - mx);
  lnZ[ib] = cnt > 0 ? lndim + mx + log(z / cnt) : -INFINITY;
  for (int p = 0; p < npar; ++p) {
    double g = 0.0;
    for (int s = 0; s < nsamp; ++s)
      if (nvec[s] > 0)
        g += exp(L[static_cast<int64_t>(s) * nb + ib] - mx) / z * dL[(static_cast<int64_t>(p) * nsamp + s) * nb + ib];
    grad[static_cast<int64_t>(p) * nb + ib] = g;
  }
}

unsigned grid_n(int64_t n) { return static_cast<unsigned>(std::min<int64_t>(kGMaxGrid, (n + kGBlock - 1) / kGBlock)); }

}  // namespace

}  // namespace ftlm_hip

using namespace ftlm_hip;

extern "C" {

int ftlm_hip_ftlm_gradient(const ftlm_op_t* op, ftlm_op_t* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                           const double* betas, int nb, double* lnZ, double* grad) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && opts != nullptr && betas != nullptr && lnZ != nullptr && nb >= 1, FTLM_EINVAL,
                 "null argument");
    FTLM_REQUIRE(npar >= 0 && npar <= 16 && (npar == 0 || (dops != nullptr && grad != nullptr)), FTLM_EINVAL,
                 "bad derivative operators");
    FTLM_REQUIRE(opts->start == FTLM_START_GAUSS_MT || opts->start == FTLM_START_PHASE, FTLM_EINVAL, "bad start");
    for (int p = 0; p < npar; ++p)
      FTLM_REQUIRE(dops[p] != nullptr && dops[p]->dim == op->dim, FTLM_EINVAL, "derivative operator of another size");
    FTLM_HIP_CHECK(hipSetDevice(op->device));
    const int64_t dim = op->dim;
    const int R = std::max(1, opts->R);
    int M = std::max(1, opts->M);
    if (opts->clamp_M_to_dim) M = static_cast<int>(std::min<int64_t>(M, std::max<int64_t>(dim, 1)));
    FTLM_REQUIRE(R <= 32, FTLM_EUNSUPPORTED, "at most 32 samples per gradient call");
    FTLM_REQUIRE(M <= 512, FTLM_EUNSUPPORTED, "at most 512 Lanczos steps per gradient call");
    if (dim <= 0) {
      for (int b = 0; b < nb; ++b) lnZ[b] = -INFINITY;
      for (int i = 0; i < npar * nb; ++i) grad[i] = 0.0;
      return;
    }
    FTLM_REQUIRE(dim > 1, FTLM_EUNSUPPORTED, "a one-state operator has the closed form -beta H00");
    const hipStream_t st = nullptr;
    const int64_t n = dim * R;
    const size_t vb = static_cast<size_t>(n);
    // buffers: q, qp, w and per parameter dq, dqp, dw, plus one scratch block
    DevBuf<double2> q(vb), qp(vb), w(vb), tmp(vb);
    std::vector<DevBuf<double2>> dq(npar), dqp(npar), dw(npar);
    for (int p = 0; p < npar; ++p) {
      dq[p].alloc(vb);
      dqp[p].alloc(vb);
      dw[p].alloc(vb);
      FTLM_HIP_CHECK(hipMemsetAsync(dq[p].p, 0, vb * sizeof(double2), st));
      FTLM_HIP_CHECK(hipMemsetAsync(dqp[p].p, 0, vb * sizeof(double2), st));
    }
    FTLM_HIP_CHECK(hipMemsetAsync(qp.p, 0, vb * sizeof(double2), st));
    // start block, normalised per sample (src/lanczos.cpp:140-175)
    std::vector<unsigned> seeds(static_cast<size_t>(R));
    for (int r = 0; r < R; ++r) seeds[r] = opts->seed + static_cast<unsigned>(r) * 100003u;
    if (opts->start == FTLM_START_PHASE) {
      DevBuf<uint32_t> ids(static_cast<size_t>(R));
      std::vector<uint32_t> hid(seeds.begin(), seeds.end());
      ids.upload(hid, st);
      hipLaunchKernelGGL(k_gphase, dim3(grid_n(dim)), dim3(kGBlock), 0, st, dim, R, ids.p, q.p);
      check_launch("k_gphase");
      FTLM_HIP_CHECK(hipStreamSynchronize(st));
    } else {
      std::vector<double2> h(vb);
      std::vector<std::complex<double>> g;
      for (int r = 0; r < R; ++r) {
        gauss_mt(dim, seeds[r], &g);
        for (int64_t i = 0; i < dim; ++i) h[static_cast<size_t>(i) * R + r] = make_double2(g[i].real(), g[i].imag());
      }
      q.upload(h, st);
    }
    const int gblk = static_cast<int>(grid_n(n));
    const int bsz = (kGBlock / R) * R;  // a multiple of R: each thread of k_gdot keeps one vector
    DevBuf<double2> part(static_cast<size_t>(gblk) * R), dres(static_cast<size_t>(R) * (2 + 2 * npar));
    std::vector<double2> hres(static_cast<size_t>(R) * (2 + 2 * npar));
    auto dots = [&](const std::vector<std::pair<const double2*, const double2*>>& pairs) {
      for (size_t i = 0; i < pairs.size(); ++i) {
        hipLaunchKernelGGL(k_gdot, dim3(gblk), dim3(bsz), 0, st, n, R, pairs[i].first, pairs[i].second, part.p);
        hipLaunchKernelGGL(k_greduce, dim3(1), dim3(64), 0, st, part.p, gblk, R, dres.p + i * R);
      }
      check_launch("k_gdot");
      FTLM_HIP_CHECK(hipMemcpyAsync(hres.data(), dres.p, pairs.size() * R * sizeof(double2), hipMemcpyDeviceToHost, st));
      FTLM_HIP_CHECK(hipStreamSynchronize(st));
    };
    // normalise the start
    {
      dots({{q.p, q.p}});
      std::vector<double> inv(static_cast<size_t>(R));
      std::vector<int> one(static_cast<size_t>(R), 1);
      std::vector<double2> hq;
      for (int r = 0; r < R; ++r) inv[r] = std::sqrt(hres[r].x);
      DevBuf<double> dinv(static_cast<size_t>(R));
      DevBuf<int> dal(static_cast<size_t>(R));
      dinv.upload(inv, st);
      dal.upload(one, st);
      for (int r = 0; r < R; ++r) FTLM_REQUIRE(inv[r] >= 1e-18, FTLM_EUNSUPPORTED, "zero start vector");
      hipLaunchKernelGGL(k_gnorm_q, dim3(grid_n(n)), dim3(kGBlock), 0, st, n, R, q.p, dinv.p, dal.p);
      check_launch("k_gnorm_q");
    }
    std::vector<double> ha(static_cast<size_t>(R) * M, 0.0), hb(static_cast<size_t>(R) * M, 0.0);
    std::vector<double> hda(static_cast<size_t>(npar) * R * M, 0.0), hdb(static_cast<size_t>(npar) * R * M, 0.0);
    std::vector<int> used(static_cast<size_t>(R), 0), alive(static_cast<size_t>(R), 1);
    std::vector<double> bprev(static_cast<size_t>(R), 0.0), dbprev(static_cast<size_t>(npar) * R, 0.0);
    DevBuf<double2> dz_dev(static_cast<size_t>(R) * (1 + npar));
    DevBuf<double> dbp(static_cast<size_t>(R) * (1 + npar)), dbeta(static_cast<size_t>(R) * (1 + npar));
    DevBuf<int> dalive(static_cast<size_t>(R));
    for (int k = 0; k < M; ++k) {
      // w = H q;  dw_p = dH_p q + H dq_p
      hv_apply(op, q.p, w.p, R, FTLM_C128, nullptr, nullptr, st);
      for (int p = 0; p < npar; ++p) {
        hv_apply(dops[p], q.p, tmp.p, R, FTLM_C128, nullptr, nullptr, st);
        hv_apply(op, dq[p].p, dw[p].p, R, FTLM_C128, nullptr, nullptr, st);
        hipLaunchKernelGGL(k_gadd, dim3(grid_n(n)), dim3(kGBlock), 0, st, n, dw[p].p, tmp.p);
      }
      check_launch("gradient H*v");
      // z = <q|w>;  dz_p = <dq_p|w> + <q|dw_p>
      std::vector<std::pair<const double2*, const double2*>> pr{{q.p, w.p}};
      for (int p = 0; p < npar; ++p) {
        pr.push_back({dq[p].p, w.p});
        pr.push_back({q.p, dw[p].p});
      }
      dots(pr);
      std::vector<double2> z(static_cast<size_t>(R) * (1 + npar));
      for (int r = 0; r < R; ++r) {
        z[r] = alive[r] ? hres[r] : make_double2(0.0, 0.0);
        if (alive[r]) ha[static_cast<size_t>(r) * M + k] = hres[r].x;
        for (int p = 0; p < npar; ++p) {
          const double2 a = hres[static_cast<size_t>(1 + 2 * p) * R + r], b = hres[static_cast<size_t>(2 + 2 * p) * R + r];
          const double2 dz = alive[r] ? make_double2(a.x + b.x, a.y + b.y) : make_double2(0.0, 0.0);
          z[static_cast<size_t>(1 + p) * R + r] = dz;
          if (alive[r]) hda[(static_cast<size_t>(p) * R + r) * M + k] = dz.x;
        }
      }
      std::vector<double> bpv(static_cast<size_t>(R) * (1 + npar));
      for (int r = 0; r < R; ++r) {
        bpv[r] = bprev[r];
        for (int p = 0; p < npar; ++p) bpv[static_cast<size_t>(1 + p) * R + r] = dbprev[static_cast<size_t>(p) * R + r];
      }
      dz_dev.upload(z, st);
      dbp.upload(bpv, st);
      dalive.upload(alive, st);
      const int use_prev = k > 0 ? 1 : 0;
      // tangents first: they read the un-updated w only through z (already taken)
      for (int p = 0; p < npar; ++p)
        hipLaunchKernelGGL(k_gupd_dw, dim3(grid_n(n)), dim3(kGBlock), 0, st, n, R, dw[p].p, q.p, dq[p].p, qp.p,
                           dqp[p].p, dz_dev.p, dz_dev.p + static_cast<size_t>(1 + p) * R, dbp.p,
                           dbp.p + static_cast<size_t>(1 + p) * R, dalive.p, use_prev);
      hipLaunchKernelGGL(k_gupd_w, dim3(grid_n(n)), dim3(kGBlock), 0, st, n, R, w.p, q.p, qp.p, dz_dev.p, dbp.p,
                         dalive.p, use_prev);
      check_launch("gradient update");
      // |w|^2 and <w|dw_p>
      std::vector<std::pair<const double2*, const double2*>> pn{{w.p, w.p}};
      for (int p = 0; p < npar; ++p) pn.push_back({w.p, dw[p].p});
      dots(pn);
      std::vector<double> bet(static_cast<size_t>(R) * (1 + npar), 0.0);
      for (int r = 0; r < R; ++r) {
        if (!alive[r]) continue;
        used[r] = k + 1;
        const double nw2 = hres[r].x;
        if (nw2 < 1e-24) {
          alive[r] = 0;
          continue;
        }
        const double b = std::sqrt(nw2);
        hb[static_cast<size_t>(r) * M + k] = b;
        bet[r] = b;
        bprev[r] = b;
        for (int p = 0; p < npar; ++p) {
          const double db = hres[static_cast<size_t>(1 + p) * R + r].x / b;
          hdb[(static_cast<size_t>(p) * R + r) * M + k] = db;
          bet[static_cast<size_t>(1 + p) * R + r] = db;
          dbprev[static_cast<size_t>(p) * R + r] = db;
        }
      }
      bool any = false;
      for (int r = 0; r < R; ++r) any = any || alive[r];
      if (!any) break;
      if (k + 1 == M) break;
      dbeta.upload(bet, st);
      dalive.upload(alive, st);
      hipLaunchKernelGGL(k_gnorm_q, dim3(grid_n(n)), dim3(kGBlock), 0, st, n, R, w.p, dbeta.p, dalive.p);
      for (int p = 0; p < npar; ++p)
        hipLaunchKernelGGL(k_gnorm_dq, dim3(grid_n(n)), dim3(kGBlock), 0, st, n, R, dw[p].p, w.p, dbeta.p,
                           dbeta.p + static_cast<size_t>(1 + p) * R, dalive.p);
      check_launch("gradient normalise");
      // rotate: (qp, q, w) <- (q, w, qp)
      std::swap(qp.p, q.p);
      std::swap(q.p, w.p);
      for (int p = 0; p < npar; ++p) {
        std::swap(dqp[p].p, dq[p].p);
        std::swap(dq[p].p, dw[p].p);
      }
    }
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
    // quadrature derivatives
    DevBuf<double> da(ha.size()), db(hb.size()), dda(std::max<size_t>(hda.size(), 1)), ddb(std::max<size_t>(hdb.size(), 1));
    da.upload(ha, st);
    db.upload(hb, st);
    if (npar) {
      dda.upload(hda, st);
      ddb.upload(hdb, st);
    }
    DevBuf<int> dn(static_cast<size_t>(R));
    dn.upload(used, st);
    DevBuf<double> dbt(static_cast<size_t>(nb));
    dbt.upload(betas, static_cast<size_t>(nb), st);
    DevBuf<double> Uw(static_cast<size_t>(R) * M * M), Wm(static_cast<size_t>(R) * M * M), lam(static_cast<size_t>(R) * 2 * M);
    DevBuf<double> Ls(static_cast<size_t>(R) * nb), dLs(static_cast<size_t>(std::max(npar, 1)) * R * nb);
    DevBuf<int> status(1);
    FTLM_HIP_CHECK(hipMemsetAsync(status.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_grad_tridiag, dim3(grid_for(R, 32)), dim3(32), 0, st, da.p, db.p, dda.p, ddb.p, dn.p, R, M,
                       npar, dbt.p, nb, Uw.p, Wm.p, lam.p, Ls.p, dLs.p, status.p);
    check_launch("k_grad_tridiag");
    DevBuf<double> dl(static_cast<size_t>(nb)), dg(static_cast<size_t>(std::max(npar, 1)) * nb);
    hipLaunchKernelGGL(k_grad_combine, dim3(grid_for(nb, 64)), dim3(64), 0, st, Ls.p, dLs.p, dn.p, R, npar, nb,
                       std::log(static_cast<double>(dim)), dl.p, dg.p);
    check_launch("k_grad_combine");
    int bad = 0;
    FTLM_HIP_CHECK(hipMemcpy(&bad, status.p, sizeof(int), hipMemcpyDeviceToHost));
    FTLM_REQUIRE(bad == 0, FTLM_EHIP, "tridiagonal QL did not converge");
    FTLM_HIP_CHECK(hipMemcpy(lnZ, dl.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
    if (npar) FTLM_HIP_CHECK(hipMemcpy(grad, dg.p, sizeof(double) * npar * nb, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
