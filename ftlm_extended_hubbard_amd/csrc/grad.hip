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
#include <map>
#include <memory>
#include <random>
#include <vector>

#include "common.hpp"

namespace ftlm_hip {

void gauss_interleaved(int64_t dim, int R, const std::vector<unsigned>& seeds, std::vector<double2>* h);

namespace {

constexpr int kGBlock = 256;
constexpr int kGMaxGrid = 4096;
constexpr int kGMaxPar = 16;
constexpr int kGMaxM = 512;

// Tangent buffers of every parameter p: dq_p, dq'_p, dw_p (= H dq_p) and tmp_p (= dH_p q).
struct GradVecs {
  double2* dq[kGMaxPar];
  double2* dqp[kGMaxPar];
  double2* dw[kGMaxPar];
  double2* tmp[kGMaxPar];
};

// Elements of the flat [dim][R] block owned by one thread all belong to vector r = tid % R: the
// block size is a multiple of R and so is the grid stride.
__device__ inline void reduce_store(double sr, double si, int R, int bsz, double* rr, double* ri, double2* out) {
  rr[threadIdx.x] = sr;
  ri[threadIdx.x] = si;
  __syncthreads();
  if (threadIdx.x < R) {
    double x = 0.0, y = 0.0;
    for (int t = threadIdx.x; t < bsz; t += R) {
      x += rr[t];
      y += ri[t];
    }
    out[threadIdx.x] = make_double2(x, y);
  }
  __syncthreads();
}

// Block partials of z_r = <q|w> and dz_{p,r} = <dq_p|w> + <q|dw_p + tmp_p>, written to
// part[(block * (1 + npar) + quantity) * R + r].
__global__ __launch_bounds__(kGBlock) void k_gstep_dots(int64_t n, int R, int npar, const double2* __restrict__ q,
                                                        const double2* __restrict__ w, GradVecs V,
                                                        double2* __restrict__ part) {
  const int bsz = (kGBlock / R) * R;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const int64_t f0 = static_cast<int>(threadIdx.x) < bsz ? static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x : n;
  double zr = 0.0, zi = 0.0, dr[kGMaxPar], di[kGMaxPar];
#pragma unroll
  for (int p = 0; p < kGMaxPar; ++p) dr[p] = di[p] = 0.0;
  for (int64_t f = f0; f < n; f += stride) {
    const double2 a = q[f], b = w[f];
    zr += a.x * b.x + a.y * b.y;
    zi += a.x * b.y - a.y * b.x;
#pragma unroll
    for (int p = 0; p < kGMaxPar; ++p)
      if (p < npar) {
        const double2 da = V.dq[p][f], dw = V.dw[p][f], t = V.tmp[p][f];
        const double2 dwt = make_double2(dw.x + t.x, dw.y + t.y);
        dr[p] += da.x * b.x + da.y * b.y + a.x * dwt.x + a.y * dwt.y;
        di[p] += da.x * b.y - da.y * b.x + a.x * dwt.y - a.y * dwt.x;
      }
  }
  __shared__ double rr[kGBlock], ri[kGBlock];
  const int Q = 1 + npar;
  double2* o = part + static_cast<int64_t>(blockIdx.x) * Q * R;
  reduce_store(zr, zi, R, bsz, rr, ri, o);
#pragma unroll
  for (int p = 0; p < kGMaxPar; ++p)
    if (p < npar) reduce_store(dr[p], di[p], R, bsz, rr, ri, o + static_cast<int64_t>(1 + p) * R);
}

// out[j] = sum over blocks of part[block * QR + j], one workgroup per output.
__global__ __launch_bounds__(kGBlock) void k_gsum(const double2* __restrict__ part, int nblk, int QR,
                                                  double2* __restrict__ out) {
  const int j = blockIdx.x;
  double x = 0.0, y = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    const double2 v = part[static_cast<int64_t>(b) * QR + j];
    x += v.x;
    y += v.y;
  }
  __shared__ double rx[kGBlock], ry[kGBlock];
  rx[threadIdx.x] = x;
  ry[threadIdx.x] = y;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (static_cast<int>(threadIdx.x) < h) {
      rx[threadIdx.x] += rx[threadIdx.x + h];
      ry[threadIdx.x] += ry[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = make_double2(rx[0], ry[0]);
}

// The three-term update of the recurrence and of its tangents, with the norms that follow:
//     w    <- w - z q - b' q'
//     dw_p <- dw_p + tmp_p - (dz_p q + z dq_p + db'_p q' + b' dq'_p)
// (the q' terms only when use_prev; dead vectors are zeroed), and block partials of |w|^2 and
// Re <w|dw_p>.  z, dz: zz[(1 + p) * R + r]; b', db': bp[(1 + p) * R + r].
__global__ __launch_bounds__(kGBlock) void k_gstep_update(int64_t n, int R, int npar, const double2* __restrict__ q,
                                                          const double2* __restrict__ qp, double2* __restrict__ w,
                                                          GradVecs V, const double2* __restrict__ zz,
                                                          const double* __restrict__ bp, const int* __restrict__ alive,
                                                          int use_prev, double2* __restrict__ part) {
  const int bsz = (kGBlock / R) * R;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const int64_t f0 = static_cast<int>(threadIdx.x) < bsz ? static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x : n;
  const int r = static_cast<int>(threadIdx.x) % R;
  const bool live = alive[r] != 0;
  const double2 z = zz[r];
  const double b1 = use_prev ? bp[r] : 0.0;
  double nrm = 0.0, cr[kGMaxPar];
#pragma unroll
  for (int p = 0; p < kGMaxPar; ++p) cr[p] = 0.0;
  for (int64_t f = f0; f < n; f += stride) {
    const double2 a = q[f];
    const double2 b = use_prev ? qp[f] : make_double2(0.0, 0.0);
    double2 c = w[f];
    c.x -= z.x * a.x - z.y * a.y + b1 * b.x;
    c.y -= z.x * a.y + z.y * a.x + b1 * b.y;
    if (!live) c = make_double2(0.0, 0.0);
    w[f] = c;
    nrm += c.x * c.x + c.y * c.y;
#pragma unroll
    for (int p = 0; p < kGMaxPar; ++p)
      if (p < npar) {
        const double2 dz = zz[static_cast<int64_t>(1 + p) * R + r];
        const double db1 = use_prev ? bp[static_cast<int64_t>(1 + p) * R + r] : 0.0;
        const double2 da = V.dq[p][f], t = V.tmp[p][f];
        const double2 dpb = use_prev ? V.dqp[p][f] : make_double2(0.0, 0.0);
        double2 d = V.dw[p][f];
        d.x += t.x - (dz.x * a.x - dz.y * a.y + z.x * da.x - z.y * da.y + db1 * b.x + b1 * dpb.x);
        d.y += t.y - (dz.x * a.y + dz.y * a.x + z.x * da.y + z.y * da.x + db1 * b.y + b1 * dpb.y);
        if (!live) d = make_double2(0.0, 0.0);
        V.dw[p][f] = d;
        cr[p] += c.x * d.x + c.y * d.y;
      }
  }
  __shared__ double rr[kGBlock], ri[kGBlock];
  const int Q = 1 + npar;
  double2* o = part + static_cast<int64_t>(blockIdx.x) * Q * R;
  reduce_store(nrm, 0.0, R, bsz, rr, ri, o);
#pragma unroll
  for (int p = 0; p < kGMaxPar; ++p)
    if (p < npar) reduce_store(cr[p], 0.0, R, bsz, rr, ri, o + static_cast<int64_t>(1 + p) * R);
}

// q_new = w / b (in place in w);  dq_new,p = (dw_p - db_p q_new) / b (in place in dw_p).
// b, db: bt[(1 + p) * R + r]; dead vectors are zeroed.
__global__ __launch_bounds__(kGBlock) void k_gstep_norm(int64_t n, int R, int npar, double2* __restrict__ w,
                                                        GradVecs V, const double* __restrict__ bt,
                                                        const int* __restrict__ alive) {
  const int bsz = (kGBlock / R) * R;
  if (static_cast<int>(threadIdx.x) >= bsz) return;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const int r = static_cast<int>(threadIdx.x) % R;
  const bool live = alive[r] != 0;
  const double inv = live ? 1.0 / bt[r] : 0.0;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x; f < n; f += stride) {
    const double2 c = w[f];
    const double2 qn = make_double2(c.x * inv, c.y * inv);
    w[f] = qn;
#pragma unroll
    for (int p = 0; p < kGMaxPar; ++p)
      if (p < npar) {
        const double db = live ? bt[static_cast<int64_t>(1 + p) * R + r] : 0.0;
        const double2 d = V.dw[p][f];
        V.dw[p][f] = make_double2((d.x - db * qn.x) * inv, (d.y - db * qn.y) * inv);
      }
  }
}

// q /= |q| per vector (the start block); norms2[r] = |q_r|^2.
__global__ __launch_bounds__(kGBlock) void k_gscale(int64_t n, int R, double2* __restrict__ q,
                                                    const double2* __restrict__ norms2) {
  const int bsz = (kGBlock / R) * R;
  if (static_cast<int>(threadIdx.x) >= bsz) return;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const double inv = 1.0 / sqrt(norms2[static_cast<int>(threadIdx.x) % R].x);
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x; f < n; f += stride) {
    const double2 c = q[f];
    q[f] = make_double2(c.x * inv, c.y * inv);
  }
}

// Converged implicit-QL eigen-decomposition of the tridiagonal (d, e) of length n by one
// workgroup: thread 0 runs the scalar recurrence in LDS and records each sweep's plane rotations in
// (rc, rs); every thread then applies the sweep to its rows of U (U[i * n + a]: component a of
// eigenvector i, so a sweep's loads are coalesced over a).  d, e, rc, rs, ctl live in LDS; U
// (n * n) in global memory.  Eigenvalues are left in d.  Returns nonzero (on every thread) when an
// eigenvalue needs more than 200 sweeps.
__device__ int ql_vectors(double* d, double* e, double* rc, double* rs, int* ctl, double* U, int n) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int64_t f = tid; f < static_cast<int64_t>(n) * n; f += nt) U[f] = (f / n == f % n) ? 1.0 : 0.0;
  __syncthreads();
  int bad = 0, iter = 0;
  for (int l = 0; l < n && !bad; ++l) {
    if (tid == 0) iter = 0;
    while (true) {
      if (tid == 0) {
        int m = l;
        for (; m < n - 1; ++m) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= 1e-300 + 2.220446049250313e-16 * dd) break;
        }
        int flag = 0, lo = m;
        if (m == l) {
          flag = 1;
        } else if (++iter > 200) {
          flag = 2;
        } else {
          double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
          double r = hypot(g, 1.0);
          g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
          double sn = 1.0, c = 1.0, p = 0.0;
          bool deflated = false;
          for (int i = m - 1; i >= l; --i) {
            const double f = sn * e[i];
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
            rc[i] = c;
            rs[i] = sn;
            lo = i;
          }
          if (!deflated) {
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
          }
        }
        ctl[0] = flag;
        ctl[1] = m;
        ctl[2] = lo;
      }
      __syncthreads();
      const int flag = ctl[0], m = ctl[1], lo = ctl[2];
      if (flag == 0 && lo < m) {
        // rotations i = m-1 .. lo of columns (i, i+1), in that order
        for (int a = tid; a < n; a += nt) {
          double carry = U[static_cast<int64_t>(m) * n + a];
          for (int i = m - 1; i >= lo; --i) {
            const double xi = U[static_cast<int64_t>(i) * n + a];
            U[static_cast<int64_t>(i + 1) * n + a] = rs[i] * xi + rc[i] * carry;
            carry = rc[i] * xi - rs[i] * carry;
          }
          U[static_cast<int64_t>(lo) * n + a] = carry;
        }
      }
      __syncthreads();
      if (flag == 2) bad = 1;
      if (flag != 0) break;
    }
  }
  return bad;
}

// One workgroup per sample s (gout, when given, receives dL_s/dalpha_k and dL_s/dbeta_k for
// every beta): eigen-decomposition of T_s (ql_vectors), then for every beta L_s(beta) and
// dL_s/dtheta_p.  alpha/beta: [s][Mmax]; dalpha/dbeta: [p][s][Mmax].  Outputs L [s][nb],
// dL [p][s][nb].
__global__ __launch_bounds__(kGBlock) void k_grad_tridiag(
    const double* __restrict__ alpha, const double* __restrict__ beta, const double* __restrict__ dalpha,
    const double* __restrict__ dbeta, const int* __restrict__ nvec, int nsamp, int Mmax, int npar,
    const double* __restrict__ betas, int nb, double* __restrict__ Uw, double* __restrict__ W,
    double* __restrict__ L, double* __restrict__ dL, double* __restrict__ gout, int* __restrict__ status) {
  const int s = blockIdx.x;
  const int n = nvec[s];
  if (n <= 0) return;  // uniform over the workgroup
  const int tid = threadIdx.x, nt = blockDim.x;
  double* U = Uw + static_cast<int64_t>(s) * Mmax * Mmax;
  double* Wm = W + static_cast<int64_t>(s) * Mmax * Mmax;
  __shared__ double d[kGMaxM], e[kGMaxM], rc[kGMaxM], rs[kGMaxM], ex[kGMaxM], g0[kGMaxM], g1[kGMaxM];
  __shared__ int ctl[3];
  __shared__ double lz_sh;
  for (int i = tid; i < n; i += nt) {
    d[i] = alpha[static_cast<int64_t>(s) * Mmax + i];
    e[i] = i < n - 1 ? beta[static_cast<int64_t>(s) * Mmax + i] : 0.0;
  }
  __syncthreads();
  const int bad = ql_vectors(d, e, rc, rs, ctl, U, n);
  if (bad) {
    if (tid == 0) atomicAdd(status, 1);
    return;  // uniform: every thread saw flag 2
  }
  for (int ib = 0; ib < nb; ++ib) {
    const double bt = betas[ib];
    // L = ln sum_i u_i^2 exp(-bt lambda_i), shifted by the largest exponent (u_i = U[i * n])
    if (tid == 0) {
      double mx = -INFINITY;
      for (int i = 0; i < n; ++i) {
        const double u = U[static_cast<int64_t>(i) * n];
        if (u != 0.0) mx = fmax(mx, -bt * d[i] + log(u * u));
      }
      double z = 0.0;
      for (int i = 0; i < n; ++i) {
        const double u = U[static_cast<int64_t>(i) * n];
        if (u != 0.0) z += exp(-bt * d[i] + log(u * u) - mx);
      }
      lz_sh = mx + log(z);
      L[static_cast<int64_t>(s) * nb + ib] = lz_sh;
    }
    __syncthreads();
    const double lz = lz_sh;
    for (int i = tid; i < n; i += nt) ex[i] = exp(-bt * d[i] - lz);
    __syncthreads();
    // W_ij = u_i u_j g[lambda_i, lambda_j] / Z, g(lambda) = exp(-bt lambda): close eigenvalues
    // through -bt exp(-bt mean) sinh(x)/x, x = bt (lambda_i - lambda_j) / 2; otherwise the quotient
    for (int64_t f = tid; f < static_cast<int64_t>(n) * n; f += nt) {
      const int i = static_cast<int>(f / n), j = static_cast<int>(f % n);
      const double ui = U[static_cast<int64_t>(i) * n], uj = U[static_cast<int64_t>(j) * n];
      double v = 0.0;
      if (ui != 0.0 && uj != 0.0) {
        const double xd = 0.5 * bt * (d[i] - d[j]);
        if (fabs(xd) < 1e-3)
          v = ui * uj * (-bt) * exp(-bt * 0.5 * (d[i] + d[j]) - lz) * (1.0 + xd * xd / 6.0);
        else
          v = ui * uj * (ex[i] - ex[j]) / (d[i] - d[j]);
      }
      Wm[f] = v;
    }
    __syncthreads();
    // G = U W U^T on the diagonal and the first off-diagonal: row k of U is (U[i * n + k])_i
    for (int k = tid; k < n; k += nt) {
      double gkk = 0.0, gk1 = 0.0;
      for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int j = 0; j < n; ++j) t += Wm[static_cast<int64_t>(i) * n + j] * U[static_cast<int64_t>(j) * n + k];
        gkk += U[static_cast<int64_t>(i) * n + k] * t;
        if (k + 1 < n) gk1 += U[static_cast<int64_t>(i) * n + k + 1] * t;
      }
      g0[k] = gkk;
      g1[k] = gk1;
      if (gout) {  // dL/dalpha_k = G_kk, dL/dbeta_k = 2 G_{k,k+1}: [beta][s][k][2]
        double* o = gout + ((static_cast<int64_t>(ib) * nsamp + s) * Mmax + k) * 2;
        o[0] = gkk;
        o[1] = k + 1 < n ? 2.0 * gk1 : 0.0;
      }
    }
    __syncthreads();
    for (int p = tid; p < npar; p += nt) {
      const int64_t o = (static_cast<int64_t>(p) * nsamp + s) * Mmax;
      double acc = 0.0;
      for (int k = 0; k < n; ++k) {
        acc += g0[k] * dalpha[o + k];
        if (k + 1 < n) acc += 2.0 * g1[k] * dbeta[o + k];
      }
      dL[(static_cast<int64_t>(p) * nsamp + s) * nb + ib] = acc;
    }
    __syncthreads();
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

// FTLM estimator of observables, per sample s and beta: with T_s = U diag(lambda) U^T, u_j = U[j][0]
// (= <r|psi_j>) and a_k = <q_k|A r> (so <psi_j|A|r> = sum_k U[j][k] a_k),
//     num_s = Re sum_j e^{-beta lambda_j} u_j <psi_j|A|r> / sum_j e^{-beta lambda_j} u_j^2,
// and L_s as in k_grad_tridiag; <A> = sum_s softmax(L)_s num_s.  One workgroup per sample.
// a: [p][s][Mmax] complex.  Outputs L [s][nb], num [p][s][nb].
__global__ __launch_bounds__(kGBlock) void k_obs_tridiag(const double* __restrict__ alpha,
                                                         const double* __restrict__ beta,
                                                         const int* __restrict__ nvec, int nsamp, int Mmax,
                                                         int nobs, const double2* __restrict__ a,
                                                         const double* __restrict__ betas, int nb,
                                                         double* __restrict__ Uw, double* __restrict__ L,
                                                         double* __restrict__ num, int* __restrict__ status) {
  const int s = blockIdx.x;
  const int n = nvec[s];
  if (n <= 0) return;  // uniform over the workgroup
  const int tid = threadIdx.x, nt = blockDim.x;
  double* U = Uw + static_cast<int64_t>(s) * Mmax * Mmax;
  __shared__ double d[kGMaxM], e[kGMaxM], rc[kGMaxM], rs[kGMaxM], pa[kGMaxM];
  __shared__ int ctl[3];
  __shared__ double red[kGBlock];
  __shared__ double lz_sh;
  for (int i = tid; i < n; i += nt) {
    d[i] = alpha[static_cast<int64_t>(s) * Mmax + i];
    e[i] = i < n - 1 ? beta[static_cast<int64_t>(s) * Mmax + i] : 0.0;
  }
  __syncthreads();
  if (ql_vectors(d, e, rc, rs, ctl, U, n)) {
    if (tid == 0) atomicAdd(status, 1);
    return;  // uniform
  }
  for (int p = 0; p < nobs; ++p) {
    // pa[j] = u_j Re <psi_j|A|r>
    const double2* ap = a + (static_cast<int64_t>(p) * nsamp + s) * Mmax;
    for (int j = tid; j < n; j += nt) {
      double v = 0.0;
      for (int k = 0; k < n; ++k) v += U[static_cast<int64_t>(j) * n + k] * ap[k].x;
      pa[j] = U[static_cast<int64_t>(j) * n] * v;
    }
    __syncthreads();
    for (int ib = 0; ib < nb; ++ib) {
      const double bt = betas[ib];
      if (tid == 0) {  // L = ln sum_j u_j^2 e^{-bt lambda_j}, shifted by the largest exponent
        double mx = -INFINITY;
        for (int i = 0; i < n; ++i) {
          const double u = U[static_cast<int64_t>(i) * n];
          if (u != 0.0) mx = fmax(mx, -bt * d[i] + log(u * u));
        }
        double z = 0.0;
        for (int i = 0; i < n; ++i) {
          const double u = U[static_cast<int64_t>(i) * n];
          if (u != 0.0) z += exp(-bt * d[i] + log(u * u) - mx);
        }
        lz_sh = mx + log(z);
        if (p == 0) L[static_cast<int64_t>(s) * nb + ib] = lz_sh;
      }
      __syncthreads();
      const double lz = lz_sh;
      double acc = 0.0;
      for (int j = tid; j < n; j += nt) acc += exp(-bt * d[j] - lz) * pa[j];
      red[tid] = acc;
      __syncthreads();
      for (int h = nt / 2; h > 0; h >>= 1) {
        if (tid < h) red[tid] += red[tid + h];
        __syncthreads();
      }
      if (tid == 0) num[(static_cast<int64_t>(p) * nsamp + s) * nb + ib] = red[0];
      __syncthreads();
    }
  }
}

// Block partials of <q | b_p> for p < np (b_p through GradVecs::tmp), [block][np][R].
__global__ __launch_bounds__(kGBlock) void k_mdot(int64_t n, int R, int np, const double2* __restrict__ q, GradVecs V,
                                                  double2* __restrict__ part) {
  const int bsz = (kGBlock / R) * R;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const int64_t f0 = static_cast<int>(threadIdx.x) < bsz ? static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x : n;
  double xr[kGMaxPar], xi[kGMaxPar];
#pragma unroll
  for (int p = 0; p < kGMaxPar; ++p) xr[p] = xi[p] = 0.0;
  for (int64_t f = f0; f < n; f += stride) {
    const double2 a = q[f];
#pragma unroll
    for (int p = 0; p < kGMaxPar; ++p)
      if (p < np) {
        const double2 b = V.tmp[p][f];
        xr[p] += a.x * b.x + a.y * b.y;
        xi[p] += a.x * b.y - a.y * b.x;
      }
  }
  __shared__ double rr[kGBlock], ri[kGBlock];
  double2* o = part + static_cast<int64_t>(blockIdx.x) * np * R;
#pragma unroll
  for (int p = 0; p < kGMaxPar; ++p)
    if (p < np) reduce_store(xr[p], xi[p], R, bsz, rr, ri, o + static_cast<int64_t>(p) * R);
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

// ------------------------------------------------------------------ adjoint (reverse-mode) sweep
// Elementwise kernels of the backward recurrence; per-vector scalars are indexed by r.

// wb = c1 A + c2 qn (into wb); partials [block][2][R]: <q|wb> and (Re <wb|qm>, 0).
__global__ __launch_bounds__(kGBlock) void k_adj_wb(int64_t n, int R, const double2* __restrict__ A,
                                                    const double2* __restrict__ qn, const double2* __restrict__ q,
                                                    const double2* __restrict__ qm, const double* __restrict__ c1,
                                                    const double* __restrict__ c2, double2* __restrict__ wb,
                                                    double2* __restrict__ part) {
  const int bsz = (kGBlock / R) * R;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const int64_t f0 = static_cast<int>(threadIdx.x) < bsz ? static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x : n;
  const int r = static_cast<int>(threadIdx.x) % R;
  const double a1 = c1[r], a2 = c2[r];
  double cr = 0.0, ci = 0.0, rq = 0.0;
  for (int64_t f = f0; f < n; f += stride) {
    const double2 x = A[f], y = qn[f], qq = q[f], pm = qm[f];
    const double2 v = make_double2(a1 * x.x + a2 * y.x, a1 * x.y + a2 * y.y);
    wb[f] = v;
    cr += qq.x * v.x + qq.y * v.y;
    ci += qq.x * v.y - qq.y * v.x;
    rq += v.x * pm.x + v.y * pm.y;
  }
  __shared__ double rr[kGBlock], ri[kGBlock];
  double2* o = part + static_cast<int64_t>(blockIdx.x) * 2 * R;
  reduce_store(cr, ci, R, bsz, rr, ri, o);
  reduce_store(rq, 0.0, R, bsz, rr, ri, o + R);
}

// wa = wb + zb q (in place).
__global__ __launch_bounds__(kGBlock) void k_adj_wa(int64_t n, int R, double2* __restrict__ wa,
                                                    const double2* __restrict__ q, const double2* __restrict__ zb) {
  const int bsz = (kGBlock / R) * R;
  if (static_cast<int>(threadIdx.x) >= bsz) return;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const double2 z = zb[static_cast<int>(threadIdx.x) % R];
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x; f < n; f += stride) {
    const double2 a = wa[f], x = q[f];
    wa[f] = make_double2(a.x + z.x * x.x - z.y * x.y, a.y + z.x * x.y + z.y * x.x);
  }
}

// With wb = wa - zb q and the forward w_a = bk X + z q + bm qm (X = q_{k+1}, or the unnormalised
// w of a sample's last step with bk = 1):
//     P <- P - conj(z) wb + conj(zb) w_a + Y      (the adjoint of q_k)
//     A <- -bm wb                                  (first part of the adjoint of q_{k-1})
// partials [block][1][R]: (Re <P|q>, 0).
__global__ __launch_bounds__(kGBlock) void k_adj_qbar(int64_t n, int R, double2* __restrict__ P, double2* __restrict__ A,
                                                      const double2* __restrict__ wa, const double2* __restrict__ Y,
                                                      const double2* __restrict__ X, const double2* __restrict__ q,
                                                      const double2* __restrict__ qm, const double2* __restrict__ zbv,
                                                      const double2* __restrict__ zv, const double* __restrict__ bkv,
                                                      const double* __restrict__ bmv, double2* __restrict__ part) {
  const int bsz = (kGBlock / R) * R;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * bsz;
  const int64_t f0 = static_cast<int>(threadIdx.x) < bsz ? static_cast<int64_t>(blockIdx.x) * bsz + threadIdx.x : n;
  const int r = static_cast<int>(threadIdx.x) % R;
  const double2 zb = zbv[r], z = zv[r];
  const double bk = bkv[r], bm = bmv[r];
  double rho = 0.0;
  for (int64_t f = f0; f < n; f += stride) {
    const double2 a = wa[f], qq = q[f], x = X[f], pm = qm[f], y = Y[f], p0 = P[f];
    const double2 wb = make_double2(a.x - (zb.x * qq.x - zb.y * qq.y), a.y - (zb.x * qq.y + zb.y * qq.x));
    const double2 fw = make_double2(bk * x.x + z.x * qq.x - z.y * qq.y + bm * pm.x,
                                    bk * x.y + z.x * qq.y + z.y * qq.x + bm * pm.y);
    double2 o;
    o.x = p0.x - (z.x * wb.x + z.y * wb.y) + (zb.x * fw.x + zb.y * fw.y) + y.x;
    o.y = p0.y - (z.x * wb.y - z.y * wb.x) + (zb.x * fw.y - zb.y * fw.x) + y.y;
    P[f] = o;
    A[f] = make_double2(-bm * wb.x, -bm * wb.y);
    rho += o.x * qq.x + o.y * qq.y;
  }
  __shared__ double rr[kGBlock], ri[kGBlock];
  reduce_store(rho, 0.0, R, bsz, rr, ri, part + static_cast<int64_t>(blockIdx.x) * R);
}

// acc[j] += sum_i Re part[i * Q + j], one workgroup per j.
__global__ __launch_bounds__(kGBlock) void k_tsum(const double2* __restrict__ part, int64_t nrow, int Q,
                                                  double* __restrict__ acc) {
  const int j = blockIdx.x;
  double x = 0.0;
  for (int64_t i = threadIdx.x; i < nrow; i += blockDim.x) x += part[i * Q + j].x;
  __shared__ double rx[kGBlock];
  rx[threadIdx.x] = x;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (static_cast<int>(threadIdx.x) < h) rx[threadIdx.x] += rx[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[j] += rx[0];
}

// Term-split bilinear forms of a full sector: for the derivative hop tables j < ntab (coefficients
// over the sector's own bond ids, [j][spin][nbp]) Re <a| dH_j |b>, then Re <a| D |b> and
// Re <a| NN |b> (D = number of doubly occupied sites, NN = the V term's pair count), all summed over
// the R vectors.  One thread per (element, vector); the hop lists are the sector's pull lists, so
// dH_j b is gathered exactly as H b is, with the derivative coefficients in place of H's.
// partials: part[block * kTsQ + quantity].
constexpr int kTsTables = 4;
constexpr int kTsQ = kTsTables + 2;

__global__ __launch_bounds__(kGBlock) void k_bil_sector(const SectorView S, const double2* __restrict__ a,
                                                        const double2* __restrict__ b, int R,
                                                        const double2* __restrict__ dc, int ntab, int nbp,
                                                        double2* __restrict__ part) {
  extern __shared__ double2 s_dyn[];
  double2* s_dc = s_dyn;                                              // [ntab][2][nbp]
  uint32_t* s_xm = reinterpret_cast<uint32_t*>(s_dyn + ntab * 2 * nbp);  // [2][nbp]
  for (int i = threadIdx.x; i < ntab * 2 * nbp; i += blockDim.x) s_dc[i] = dc[i];
  for (int i = threadIdx.x; i < nbp; i += blockDim.x) {
    s_xm[i] = i < S.nbond_up ? S.xmask_up[i] : 0u;
    s_xm[nbp + i] = i < S.nbond_dn ? S.xmask_dn[i] : 0u;
  }
  __syncthreads();
  double acc[kTsQ];
#pragma unroll
  for (int j = 0; j < kTsQ; ++j) acc[j] = 0.0;
  const int64_t n = S.dim * R;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t e = f / R;
    const int r = static_cast<int>(f - e * R);
    const int iu = static_cast<int>(e / S.Ddn);
    const int id = static_cast<int>(e - static_cast<int64_t>(iu) * S.Ddn);
    const uint32_t u = S.up_mask[iu], d = S.dn_mask[id];
    const double2 av = a[f], bv = b[f];
    const double re = av.x * bv.x + av.y * bv.y;
    const uint4 pl = S.planes[iu];
    const int cross = __popc(d & pl.x) + 2 * __popc(d & pl.y) + 4 * __popc(d & pl.z) + 8 * __popc(d & pl.w);
    acc[kTsTables] += static_cast<double>(__popc(u & d)) * re;
    acc[kTsTables + 1] += static_cast<double>(S.nuu[iu] + S.ndd[id] + cross) * re;
    double tr[kTsTables], ti[kTsTables];
#pragma unroll
    for (int j = 0; j < kTsTables; ++j) tr[j] = ti[j] = 0.0;
    const int64_t rowbase = static_cast<int64_t>(iu) * S.Ddn;
    constexpr int G = 8;  // hops whose loads are issued together
    // spin 1 (down bonds, sources in row iu), then spin 0 (up bonds, sources in column id)
#pragma unroll
    for (int sp = 1; sp >= 0; --sp) {
      const int k0 = sp ? S.off_dn[id] : S.off_up[iu], k1 = sp ? S.off_dn[id + 1] : S.off_up[iu + 1];
      const uint32_t* E = sp ? S.ent_dn : S.ent_up;
      const uint32_t other = sp ? u : d;
      for (int k = k0; k < k1; k += G) {
        uint32_t en[G];
        double2 src[G];
#pragma unroll
        for (int j = 0; j < G; ++j) en[j] = E[min(k + j, k1 - 1)];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int64_t t = en[j] & kTargetMask;
          src[j] = b[(sp ? rowbase + t : t * S.Ddn + id) * R + r];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          if (k + j >= k1) continue;
          const uint32_t bd = (en[j] >> kTargetBits) & kBondMask;
          const bool neg = ((((en[j] >> kParityBit) & 1u) ^ (__popc(other & s_xm[sp * nbp + bd]) & 1u)) != 0);
          const double2 x = neg ? make_double2(-src[j].x, -src[j].y) : src[j];
#pragma unroll
          for (int q = 0; q < kTsTables; ++q)
            if (q < ntab) {
              const double2 c = s_dc[(q * 2 + sp) * nbp + bd];
              tr[q] += c.x * x.x - c.y * x.y;
              ti[q] += c.x * x.y + c.y * x.x;
            }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kTsTables; ++j) acc[j] += av.x * tr[j] + av.y * ti[j];
  }
  __shared__ double red[kGBlock / 64][kTsQ];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kTsQ; ++j) {
    double v = acc[j];
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wv][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < kTsQ) {
    double v = 0.0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) v += red[w][threadIdx.x];
    part[static_cast<int64_t>(blockIdx.x) * kTsQ + threadIdx.x] = make_double2(v, 0.0);
  }
}

unsigned grid_n(int64_t n) { return static_cast<unsigned>(std::min<int64_t>(kGMaxGrid, (n + kGBlock - 1) / kGBlock)); }


// Derivative operators as term-split tables over a full sector's own bonds: every derivative
// operator p is a hop part (table tab[p]) plus dU[p] D + dV[p] NN.  Possible when H is a full
// sector, every derivative operator is a sector of the same (n_sites, n_up, n_dn) whose bonds are
// bonds of H, at most kTsTables of them have hops, and V terms use H's NN pair list.
struct TermSplit {
  bool fused = false;   // a term-split form exists (full sector or momentum block)
  bool kblock = false;  // the momentum-block form (kb) rather than the sector tables
  int ntab = 0, nbp = 1;
  std::vector<int> tab;
  std::vector<double> dU, dV;
  DevBuf<double2> dc;  // [ntab][spin][nbp]
  KbTermSplit kb;
};

bool build_termsplit(const ftlm_op* op, ftlm_op* const* dops, int npar, TermSplit* T) {
  if (op->kind == kKBlock) {
    if (!kblock_termsplit(op, dops, npar, &T->kb)) return false;
    T->kblock = T->fused = true;
    return true;
  }
  if (op->kind != kSector) return false;
  const Sector* H = static_cast<const Sector*>(op);
  std::map<std::pair<int, int>, int> id[2];
  for (int sp = 0; sp < 2; ++sp)
    for (size_t b = 0; b < H->bond_key[sp].size(); ++b) id[sp][H->bond_key[sp][b]] = static_cast<int>(b);
  const int nbp = std::max(1, std::max(H->nbond[0], H->nbond[1]));
  std::vector<std::complex<double>> tabs;
  std::vector<int> tab(static_cast<size_t>(npar), -1);
  std::vector<double> dU(static_cast<size_t>(npar), 0.0), dV(static_cast<size_t>(npar), 0.0);
  int ntab = 0;
  for (int p = 0; p < npar; ++p) {
    if (dops[p]->kind != kSector) return false;
    const Sector* D = static_cast<const Sector*>(dops[p]);
    if (D->n_sites != H->n_sites || D->n_up != H->n_up || D->n_dn != H->n_dn) return false;
    if (D->nbond[0] + D->nbond[1] > 0) {
      if (ntab == kTsTables) return false;
      const int j = ntab++;
      tabs.resize(static_cast<size_t>(ntab) * 2 * nbp);
      for (int sp = 0; sp < 2; ++sp)
        for (size_t b = 0; b < D->bond_key[sp].size(); ++b) {
          const auto it = id[sp].find(D->bond_key[sp][b]);
          if (it == id[sp].end()) return false;
          tabs[(static_cast<size_t>(j) * 2 + sp) * nbp + it->second] += D->bond_coef[sp][b];
        }
      tab[p] = j;
    }
    dU[p] = D->U;
    dV[p] = D->V;
    if (D->V != 0.0 && D->pair_list != H->pair_list) return false;
  }
  const size_t lds = static_cast<size_t>(ntab) * 2 * nbp * sizeof(double2) + 2 * static_cast<size_t>(nbp) * 4;
  if (lds > 64 * 1024) return false;
  std::vector<double2> h(std::max<size_t>(tabs.size(), 1), make_double2(0.0, 0.0));
  for (size_t i = 0; i < tabs.size(); ++i) h[i] = make_double2(tabs[i].real(), tabs[i].imag());
  T->dc.alloc(h.size());
  T->dc.upload(h);
  T->ntab = ntab;
  T->nbp = nbp;
  T->tab = tab;
  T->dU = dU;
  T->dV = dV;
  T->fused = true;
  return true;
}

}  // namespace

}  // namespace ftlm_hip

using namespace ftlm_hip;

// The saved state of one forward Lanczos pass for repeated adjoint sweeps.
struct ftlm_tape {
  const ftlm_op* op = nullptr;
  std::vector<ftlm_op*> dops;
  ftlm_hip::TermSplit ts;
  int R = 0, M = 0, s = 1, nb = 0, kmax = 0;
  int64_t dim = 0;
  std::vector<ftlm_hip::DevBuf<double2>> ck;  // ck[2c], ck[2c+1] = q_{cs-1}, q_{cs}
  std::vector<double2> zs;                    // [M][R] z_k (dead vectors 0)
  std::vector<double> bs;                     // [M][R] beta_k
  std::vector<int> used;                      // steps taken per vector
  std::vector<double> L, G;                   // L [R][nb]; dL/d(alpha_k, beta_k): [nb][R][M][2]
};

namespace ftlm_hip {
namespace {

// Forward pass of the adjoint method: the recurrence of ftlm_hip_ftlm_gradient without tangents,
// keeping (q_{k-1}, q_k) every s steps, then the quadrature and its derivatives per sample.
// Optional observables of the forward pass: a[(p * M + k) * R + r] = <q_k|A_p q_0> per step.
struct ObsAcc {
  std::vector<const ftlm_op*> aops;
  std::vector<double2> a;
};

void tape_forward(ftlm_tape* T, const ftlm_op* op, ftlm_op* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                  const double* betas, int nb, int ckpt, int mode, double* lnZ, ObsAcc* obs = nullptr) {
  FTLM_REQUIRE(op != nullptr && opts != nullptr && betas != nullptr && lnZ != nullptr && nb >= 1, FTLM_EINVAL,
               "null argument");
  FTLM_REQUIRE(npar >= 0 && npar <= kGMaxPar && (npar == 0 || dops != nullptr), FTLM_EINVAL,
               "bad derivative operators");
  FTLM_REQUIRE(mode == FTLM_TS_AUTO || mode == FTLM_TS_GENERIC || mode == FTLM_TS_FUSED, FTLM_EINVAL, "bad mode");
  FTLM_REQUIRE(ckpt >= 0, FTLM_EINVAL, "negative checkpoint interval");
  FTLM_REQUIRE(opts->start == FTLM_START_GAUSS_MT || opts->start == FTLM_START_PHASE, FTLM_EINVAL, "bad start");
  FTLM_REQUIRE(opts->reorth == FTLM_REORTH_NONE, FTLM_EUNSUPPORTED,
               "derivatives follow the plain (unreorthogonalised) recurrence");
  for (int p = 0; p < npar; ++p)
    FTLM_REQUIRE(dops[p] != nullptr && dops[p]->dim == op->dim, FTLM_EINVAL, "derivative operator of another size");
  FTLM_HIP_CHECK(hipSetDevice(op->device));
  const int64_t dim = op->dim;
  const int R = std::max(1, opts->R);
  int M = std::max(1, opts->M);
  if (opts->clamp_M_to_dim) M = static_cast<int>(std::min<int64_t>(M, std::max<int64_t>(dim, 1)));
  FTLM_REQUIRE(R <= 32, FTLM_EUNSUPPORTED, "at most 32 samples per gradient call");
  FTLM_REQUIRE(M <= kGMaxM, FTLM_EUNSUPPORTED, "at most 512 Lanczos steps per gradient call");
  FTLM_REQUIRE(dim > 1, FTLM_EUNSUPPORTED, "the adjoint sweep needs dim > 1 (a one-state operator has lnZ = -beta H00)");
  // the adjoint sweep applies H in place of H^dag
  FTLM_REQUIRE(op->kind != kKBlock || kblock_hermitian(op), FTLM_EUNSUPPORTED,
               "this momentum-block operator is not Hermitian (its hop coefficients are not translation "
               "invariant, e.g. a Peierls flux on the wrapped bonds); use ftlm_hip_ftlm_gradient");
  if (mode != FTLM_TS_GENERIC && npar > 0) build_termsplit(op, dops, npar, &T->ts);
  FTLM_REQUIRE(mode != FTLM_TS_FUSED || T->ts.fused || npar == 0, FTLM_EUNSUPPORTED,
               "these derivative operators have no term-split form over this operator");
  T->op = op;
  T->dops.assign(dops, dops + npar);
  T->R = R;
  T->M = M;
  T->nb = nb;
  T->dim = dim;
  const int s = std::min(M, ckpt > 0 ? ckpt : std::max(1, static_cast<int>(std::lround(std::sqrt(2.0 * M)))));
  T->s = s;
  const int nck = (M + s - 1) / s;
  const hipStream_t st = nullptr;
  const int64_t n = dim * R;
  const size_t vb = static_cast<size_t>(n);
  T->ck.resize(2 * static_cast<size_t>(nck));
  for (auto& b : T->ck) b.alloc(vb);
  DevBuf<double2> b0(vb), b1(vb), b2(vb);
  double2* qp = b0.p;
  double2* q = b1.p;
  double2* w = b2.p;
  const int bsz = (kGBlock / R) * R;
  const int gblk = static_cast<int>(std::min<int64_t>(kGMaxGrid, (n + bsz - 1) / bsz));
  DevBuf<double2> part(static_cast<size_t>(gblk) * R), dres(static_cast<size_t>(R));
  std::vector<double2> hres(static_cast<size_t>(R));
  auto collect = [&] {
    hipLaunchKernelGGL(k_gsum, dim3(R), dim3(kGBlock), 0, st, part.p, gblk, R, dres.p);
    check_launch("k_gsum");
    FTLM_HIP_CHECK(hipMemcpyAsync(hres.data(), dres.p, sizeof(double2) * R, hipMemcpyDeviceToHost, st));
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
  };
  FTLM_HIP_CHECK(hipMemsetAsync(qp, 0, vb * sizeof(double2), st));
  std::vector<unsigned> seeds(static_cast<size_t>(R));
  for (int r = 0; r < R; ++r) seeds[r] = opts->seed + static_cast<unsigned>(r) * 100003u;
  if (opts->start == FTLM_START_PHASE) {
    DevBuf<uint32_t> ids(static_cast<size_t>(R));
    std::vector<uint32_t> hid(seeds.begin(), seeds.end());
    ids.upload(hid, st);
    hipLaunchKernelGGL(k_gphase, dim3(grid_n(dim)), dim3(kGBlock), 0, st, dim, R, ids.p, q);
    check_launch("k_gphase");
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
  } else {
    std::vector<double2> h;
    gauss_interleaved(dim, R, seeds, &h);
    FTLM_HIP_CHECK(hipMemcpy(q, h.data(), vb * sizeof(double2), hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_gstep_dots, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, q, q, GradVecs{}, part.p);
  check_launch("k_gstep_dots");
  collect();
  for (int r = 0; r < R; ++r) FTLM_REQUIRE(std::sqrt(hres[r].x) >= 1e-18, FTLM_EUNSUPPORTED, "zero start vector");
  hipLaunchKernelGGL(k_gscale, dim3(gblk), dim3(kGBlock), 0, st, n, R, q, dres.p);
  check_launch("k_gscale");
  // observables: b_p = A_p q_0, then <q_k|b_p> at every step
  const int nobs = obs ? static_cast<int>(obs->aops.size()) : 0;
  std::vector<DevBuf<double2>> bvec(static_cast<size_t>(nobs));
  GradVecs BV{};
  DevBuf<double2> opart(nobs ? static_cast<size_t>(gblk) * nobs * R : 0), ores(static_cast<size_t>(nobs) * R);
  std::vector<double2> hobs(static_cast<size_t>(nobs) * R);
  if (nobs) {
    obs->a.assign(static_cast<size_t>(nobs) * M * R, make_double2(0.0, 0.0));
    for (int p = 0; p < nobs; ++p) {
      bvec[p].alloc(vb);
      hv_apply(obs->aops[p], q, bvec[p].p, R, FTLM_C128, nullptr, nullptr, st);
      BV.tmp[p] = bvec[p].p;
    }
    check_launch("observable start");
  }
  T->zs.assign(static_cast<size_t>(M) * R, make_double2(0.0, 0.0));
  T->bs.assign(static_cast<size_t>(M) * R, 0.0);
  T->used.assign(static_cast<size_t>(R), 0);
  std::vector<int> alive(static_cast<size_t>(R), 1);
  std::vector<double> bprev(static_cast<size_t>(R), 0.0), bet(static_cast<size_t>(R), 0.0);
  std::vector<double2> z(static_cast<size_t>(R));
  DevBuf<double2> dz(static_cast<size_t>(R));
  DevBuf<double> dbp(static_cast<size_t>(R)), dbt(static_cast<size_t>(R));
  DevBuf<int> dalive(static_cast<size_t>(R));
  for (int k = 0; k < M; ++k) {
    if (k % s == 0) {
      const int c = k / s;
      FTLM_HIP_CHECK(hipMemcpyAsync(T->ck[2 * c].p, qp, vb * sizeof(double2), hipMemcpyDeviceToDevice, st));
      FTLM_HIP_CHECK(hipMemcpyAsync(T->ck[2 * c + 1].p, q, vb * sizeof(double2), hipMemcpyDeviceToDevice, st));
    }
    if (nobs) {
      hipLaunchKernelGGL(k_mdot, dim3(gblk), dim3(kGBlock), 0, st, n, R, nobs, q, BV, opart.p);
      hipLaunchKernelGGL(k_gsum, dim3(nobs * R), dim3(kGBlock), 0, st, opart.p, gblk, nobs * R, ores.p);
      check_launch("k_mdot");
      FTLM_HIP_CHECK(hipMemcpyAsync(hobs.data(), ores.p, sizeof(double2) * nobs * R, hipMemcpyDeviceToHost, st));
      FTLM_HIP_CHECK(hipStreamSynchronize(st));
      for (int p = 0; p < nobs; ++p)
        for (int r = 0; r < R; ++r)
          obs->a[(static_cast<size_t>(p) * M + k) * R + r] =
              alive[r] ? hobs[static_cast<size_t>(p) * R + r] : make_double2(0.0, 0.0);
    }
    hv_apply(op, q, w, R, FTLM_C128, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_gstep_dots, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, q, w, GradVecs{}, part.p);
    check_launch("k_gstep_dots");
    collect();
    for (int r = 0; r < R; ++r) {
      z[r] = alive[r] ? hres[r] : make_double2(0.0, 0.0);
      T->zs[static_cast<size_t>(k) * R + r] = z[r];
    }
    dz.upload(z, st);
    dbp.upload(bprev, st);
    dalive.upload(alive, st);
    hipLaunchKernelGGL(k_gstep_update, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, q, qp, w, GradVecs{}, dz.p, dbp.p,
                       dalive.p, k > 0 ? 1 : 0, part.p);
    check_launch("k_gstep_update");
    collect();
    std::fill(bet.begin(), bet.end(), 0.0);
    for (int r = 0; r < R; ++r) {
      if (!alive[r]) continue;
      T->used[r] = k + 1;
      const double nw2 = hres[r].x;
      if (nw2 < 1e-24) {
        alive[r] = 0;
        continue;
      }
      const double b = std::sqrt(nw2);
      T->bs[static_cast<size_t>(k) * R + r] = b;
      bet[r] = b;
      bprev[r] = b;
    }
    bool any = false;
    for (int r = 0; r < R; ++r) any = any || alive[r];
    if (!any || k + 1 == M) break;
    dbt.upload(bet, st);
    dalive.upload(alive, st);
    hipLaunchKernelGGL(k_gstep_norm, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, w, GradVecs{}, dbt.p, dalive.p);
    check_launch("k_gstep_norm");
    double2* t = qp;
    qp = q;
    q = w;
    w = t;
  }
  T->kmax = 0;
  for (int r = 0; r < R; ++r) T->kmax = std::max(T->kmax, T->used[r]);
  // quadrature: L_r(beta_b) and dL_r/d(alpha_k, beta_k)
  std::vector<double> ha(static_cast<size_t>(R) * M, 0.0), hb(static_cast<size_t>(R) * M, 0.0);
  for (int r = 0; r < R; ++r)
    for (int k = 0; k < T->used[r]; ++k) {
      ha[static_cast<size_t>(r) * M + k] = T->zs[static_cast<size_t>(k) * R + r].x;
      hb[static_cast<size_t>(r) * M + k] = T->bs[static_cast<size_t>(k) * R + r];
    }
  DevBuf<double> da(ha.size()), db(hb.size()), dbeta(static_cast<size_t>(nb));
  da.upload(ha, st);
  db.upload(hb, st);
  dbeta.upload(betas, static_cast<size_t>(nb), st);
  DevBuf<int> dn(static_cast<size_t>(R));
  dn.upload(T->used, st);
  DevBuf<double> Uw(static_cast<size_t>(R) * M * M), Wm(static_cast<size_t>(R) * M * M);
  DevBuf<double> Ls(static_cast<size_t>(R) * nb), G(static_cast<size_t>(nb) * R * M * 2);
  DevBuf<int> status(1);
  FTLM_HIP_CHECK(hipMemsetAsync(status.p, 0, sizeof(int), st));
  FTLM_HIP_CHECK(hipMemsetAsync(G.p, 0, G.n * sizeof(double), st));
  hipLaunchKernelGGL(k_grad_tridiag, dim3(R), dim3(kGBlock), 0, st, da.p, db.p, static_cast<const double*>(nullptr),
                     static_cast<const double*>(nullptr), dn.p, R, M, 0, dbeta.p, nb, Uw.p, Wm.p, Ls.p,
                     static_cast<double*>(nullptr), G.p, status.p);
  check_launch("k_grad_tridiag");
  int bad = 0;
  FTLM_HIP_CHECK(hipMemcpy(&bad, status.p, sizeof(int), hipMemcpyDeviceToHost));
  FTLM_REQUIRE(bad == 0, FTLM_EHIP, "tridiagonal QL did not converge");
  T->L.resize(Ls.n);
  T->G.resize(G.n);
  FTLM_HIP_CHECK(hipMemcpy(T->L.data(), Ls.p, T->L.size() * sizeof(double), hipMemcpyDeviceToHost));
  FTLM_HIP_CHECK(hipMemcpy(T->G.data(), G.p, T->G.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int b = 0; b < nb; ++b) {  // lnZ_b = ln dim + log-mean-exp_r L_r(b)
    double mx = -INFINITY;
    for (int r = 0; r < R; ++r) mx = std::max(mx, T->L[static_cast<size_t>(r) * nb + b]);
    double zsum = 0.0;
    for (int r = 0; r < R; ++r) zsum += std::exp(T->L[static_cast<size_t>(r) * nb + b] - mx);
    lnZ[b] = std::log(static_cast<double>(dim)) + mx + std::log(zsum / R);
  }
}

// Adjoint sweep for the objective sum_b gw[b] lnZ(beta_b): grad[p] = sum_k Re <w-bar_k| dH_p |q_k>.
void tape_backward(const ftlm_tape* T, const double* gw, double* grad) {
  const int npar = static_cast<int>(T->dops.size());
  FTLM_REQUIRE(gw != nullptr && (npar == 0 || grad != nullptr), FTLM_EINVAL, "null argument");
  if (npar == 0) return;
  const ftlm_op* op = T->op;
  FTLM_HIP_CHECK(hipSetDevice(op->device));
  const int R = T->R, M = T->M, s = T->s, nb = T->nb, kmax = T->kmax;
  const std::vector<int>& used = T->used;
  const std::vector<double>& bs = T->bs;
  const std::vector<double2>& zs = T->zs;
  const TermSplit& ts = T->ts;
  const hipStream_t st = nullptr;
  const int64_t n = T->dim * R;
  const size_t vb = static_cast<size_t>(n);
  // d(sum_b gw_b lnZ_b)/dL_r(b) = gw_b softmax_r(L(b)): adjoints of alpha_k (abar) and beta_k (bbar)
  std::vector<double> abar(static_cast<size_t>(R) * M, 0.0), bbar(static_cast<size_t>(R) * M, 0.0);
  for (int b = 0; b < nb; ++b) {
    double mx = -INFINITY;
    for (int r = 0; r < R; ++r) mx = std::max(mx, T->L[static_cast<size_t>(r) * nb + b]);
    double zsum = 0.0;
    for (int r = 0; r < R; ++r) zsum += std::exp(T->L[static_cast<size_t>(r) * nb + b] - mx);
    for (int r = 0; r < R; ++r) {
      const double wgt = gw[b] * std::exp(T->L[static_cast<size_t>(r) * nb + b] - mx) / zsum;
      for (int k = 0; k < used[r]; ++k) {
        const double* g = &T->G[((static_cast<size_t>(b) * R + r) * M + k) * 2];
        abar[static_cast<size_t>(r) * M + k] += wgt * g[0];
        bbar[static_cast<size_t>(r) * M + k] += wgt * g[1];
      }
    }
  }
  // seg[i] = q_{j0-1+i} of the segment being reversed
  std::vector<DevBuf<double2>> seg(static_cast<size_t>(s) + 2);
  for (auto& b : seg) b.alloc(vb);
  DevBuf<double2> A(vb), P(vb), WA(vb), Y(vb), tmp(ts.fused ? 0 : vb);
  const int bsz = (kGBlock / R) * R;
  const int gblk = static_cast<int>(std::min<int64_t>(kGMaxGrid, (n + bsz - 1) / bsz));
  const int gbil = ts.kblock ? kblock_bilinear_blocks(op, R)
                             : static_cast<int>(std::min<int64_t>(kGMaxGrid, (n + kGBlock - 1) / kGBlock));
  constexpr int kKbQ = 8;  // the momentum-block form's partial sums per block (4 tables + 4 diagonals)
  const int nq_bil = ts.kblock ? kKbQ : kTsQ;
  DevBuf<double2> part(std::max(static_cast<size_t>(gblk) * 2 * R, static_cast<size_t>(gbil) * nq_bil));
  DevBuf<double2> dres(static_cast<size_t>(2) * R);
  std::vector<double2> hres(static_cast<size_t>(2) * R);
  auto collect = [&](int nq) {
    hipLaunchKernelGGL(k_gsum, dim3(nq * R), dim3(kGBlock), 0, st, part.p, gblk, nq * R, dres.p);
    check_launch("k_gsum");
    FTLM_HIP_CHECK(hipMemcpyAsync(hres.data(), dres.p, sizeof(double2) * nq * R, hipMemcpyDeviceToHost, st));
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
  };
  const int nacc = ts.fused ? nq_bil : npar;
  DevBuf<double> tacc(static_cast<size_t>(nacc));
  FTLM_HIP_CHECK(hipMemsetAsync(tacc.p, 0, tacc.n * sizeof(double), st));
  FTLM_HIP_CHECK(hipMemsetAsync(A.p, 0, vb * sizeof(double2), st));
  FTLM_HIP_CHECK(hipMemsetAsync(P.p, 0, vb * sizeof(double2), st));
  std::vector<double> rho(static_cast<size_t>(R), 0.0), bacc(static_cast<size_t>(R), 0.0);
  std::vector<double> c1(static_cast<size_t>(R)), c2(static_cast<size_t>(R)), bk(static_cast<size_t>(R)),
      bm(static_cast<size_t>(R)), scale(static_cast<size_t>(R));
  std::vector<double2> zk(static_cast<size_t>(R)), zb(static_cast<size_t>(R));
  std::vector<int> live(static_cast<size_t>(R));
  DevBuf<double> d_c1(static_cast<size_t>(R)), d_c2(static_cast<size_t>(R)), d_bk(static_cast<size_t>(R)),
      d_bm(static_cast<size_t>(R)), dbt(static_cast<size_t>(R));
  DevBuf<double2> d_zb(static_cast<size_t>(R)), d_zk(static_cast<size_t>(R));
  DevBuf<int> dalive(static_cast<size_t>(R));
  const Sector* Hs = ts.fused && !ts.kblock ? static_cast<const Sector*>(op) : nullptr;
  const size_t bil_lds =
      Hs ? static_cast<size_t>(ts.ntab) * 2 * ts.nbp * sizeof(double2) + 2 * static_cast<size_t>(ts.nbp) * 4 : 0;
  auto beta_at = [&](int k, int r) { return k >= 0 ? bs[static_cast<size_t>(k) * R + r] : 0.0; };
  for (int c = (kmax - 1) / s; c >= 0; --c) {
    const int j0 = c * s, j1 = std::min(j0 + s, kmax);
    FTLM_HIP_CHECK(hipMemcpyAsync(seg[0].p, T->ck[2 * c].p, vb * sizeof(double2), hipMemcpyDeviceToDevice, st));
    FTLM_HIP_CHECK(hipMemcpyAsync(seg[1].p, T->ck[2 * c + 1].p, vb * sizeof(double2), hipMemcpyDeviceToDevice, st));
    // recompute q_{j0+1} .. q_{j1} with the stored scalars and the forward pass's kernels; at a
    // vector's last step the unnormalised w is kept
    for (int j = j0; j < j1; ++j) {
      double2* X = seg[static_cast<size_t>(j - j0 + 2)].p;
      const double2* qj = seg[static_cast<size_t>(j - j0 + 1)].p;
      const double2* qjm = seg[static_cast<size_t>(j - j0)].p;
      hv_apply(op, qj, X, R, FTLM_C128, nullptr, nullptr, st);
      for (int r = 0; r < R; ++r) {
        live[r] = j < used[r] ? 1 : 0;
        zk[r] = zs[static_cast<size_t>(j) * R + r];
        bm[r] = beta_at(j - 1, r);
        scale[r] = j + 1 < used[r] ? bs[static_cast<size_t>(j) * R + r] : 1.0;
      }
      d_zk.upload(zk, st);
      d_bm.upload(bm, st);
      dalive.upload(live, st);
      hipLaunchKernelGGL(k_gstep_update, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, qj, qjm, X, GradVecs{}, d_zk.p,
                         d_bm.p, dalive.p, j > 0 ? 1 : 0, part.p);
      dbt.upload(scale, st);
      hipLaunchKernelGGL(k_gstep_norm, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, X, GradVecs{}, dbt.p, dalive.p);
      check_launch("recompute");
    }
    for (int k = j1 - 1; k >= j0; --k) {
      const double2* qm = seg[static_cast<size_t>(k - j0)].p;
      const double2* qk = seg[static_cast<size_t>(k - j0 + 1)].p;
      const double2* X = seg[static_cast<size_t>(k - j0 + 2)].p;
      for (int r = 0; r < R; ++r) {
        const bool inner = k + 1 < used[r];  // q_{k+1} and beta_k feed the quadrature
        const double b = bs[static_cast<size_t>(k) * R + r];
        c1[r] = inner ? 1.0 / b : 0.0;
        c2[r] = inner ? bbar[static_cast<size_t>(r) * M + k] + bacc[r] - rho[r] / b : 0.0;
      }
      d_c1.upload(c1, st);
      d_c2.upload(c2, st);
      hipLaunchKernelGGL(k_adj_wb, dim3(gblk), dim3(kGBlock), 0, st, n, R, A.p, X, qk, qm, d_c1.p, d_c2.p, WA.p,
                         part.p);
      check_launch("k_adj_wb");
      collect(2);
      for (int r = 0; r < R; ++r) {
        const double2 cc = hres[r];
        zb[r] = k < used[r] ? make_double2(abar[static_cast<size_t>(r) * M + k] - cc.x, -cc.y) : make_double2(0.0, 0.0);
        bacc[r] = -hres[static_cast<size_t>(R) + r].x;  // adjoint of beta_{k-1} from this step
      }
      d_zb.upload(zb, st);
      hipLaunchKernelGGL(k_adj_wa, dim3(gblk), dim3(kGBlock), 0, st, n, R, WA.p, qk, d_zb.p);
      check_launch("k_adj_wa");
      // parameter adjoints: Re <w_a-bar | dH_p q_k>
      if (ts.kblock) {
        kblock_bilinear(op, ts.kb, WA.p, qk, R, part.p, st);
        hipLaunchKernelGGL(k_tsum, dim3(kKbQ), dim3(kGBlock), 0, st, part.p, static_cast<int64_t>(gbil), kKbQ, tacc.p);
        check_launch("k_tsum");
      } else if (ts.fused) {
        hipLaunchKernelGGL(k_bil_sector, dim3(gbil), dim3(kGBlock), bil_lds, st, Hs->view(), WA.p, qk, R, ts.dc.p,
                           ts.ntab, ts.nbp, part.p);
        hipLaunchKernelGGL(k_tsum, dim3(kTsQ), dim3(kGBlock), 0, st, part.p, static_cast<int64_t>(gbil), kTsQ, tacc.p);
        check_launch("k_bil_sector");
      } else {
        for (int p = 0; p < npar; ++p) {
          hv_apply(T->dops[p], qk, tmp.p, R, FTLM_C128, nullptr, nullptr, st);
          hipLaunchKernelGGL(k_gstep_dots, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, WA.p, tmp.p, GradVecs{},
                             part.p);
          hipLaunchKernelGGL(k_tsum, dim3(1), dim3(kGBlock), 0, st, part.p, static_cast<int64_t>(gblk) * R, 1,
                             tacc.p + p);
        }
        check_launch("derivative dots");
      }
      if (k == 0) break;  // the start vector does not depend on the parameters
      hv_apply(op, WA.p, Y.p, R, FTLM_C128, nullptr, nullptr, st);
      for (int r = 0; r < R; ++r) {
        zk[r] = zs[static_cast<size_t>(k) * R + r];
        bk[r] = k + 1 < used[r] ? bs[static_cast<size_t>(k) * R + r] : (k < used[r] ? 1.0 : 0.0);
        bm[r] = beta_at(k - 1, r);
      }
      d_zk.upload(zk, st);
      d_bk.upload(bk, st);
      d_bm.upload(bm, st);
      hipLaunchKernelGGL(k_adj_qbar, dim3(gblk), dim3(kGBlock), 0, st, n, R, P.p, A.p, WA.p, Y.p, X, qk, qm, d_zb.p,
                         d_zk.p, d_bk.p, d_bm.p, part.p);
      check_launch("k_adj_qbar");
      collect(1);
      for (int r = 0; r < R; ++r) rho[r] = hres[r].x;
      std::swap(A.p, P.p);
    }
  }
  std::vector<double> hacc(static_cast<size_t>(nacc));
  FTLM_HIP_CHECK(hipMemcpy(hacc.data(), tacc.p, hacc.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int p = 0; p < npar; ++p) {
    if (!ts.fused) {
      grad[p] = hacc[p];
      continue;
    }
    if (ts.kblock) {
      double g = 0.0;
      if (ts.kb.tab[p] >= 0) g += hacc[ts.kb.tab[p]];
      if (ts.kb.dg[p] >= 0) g += hacc[4 + ts.kb.dg[p]];
      grad[p] = g;
      continue;
    }
    double g = ts.dU[p] * hacc[kTsTables] + ts.dV[p] * hacc[kTsTables + 1];
    if (ts.tab[p] >= 0) g += hacc[ts.tab[p]];
    grad[p] = g;
  }
}

}  // namespace
}  // namespace ftlm_hip

extern "C" {

int ftlm_hip_ftlm_gradient(const ftlm_op_t* op, ftlm_op_t* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                           const double* betas, int nb, double* lnZ, double* grad) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && opts != nullptr && betas != nullptr && lnZ != nullptr && nb >= 1, FTLM_EINVAL,
                 "null argument");
    FTLM_REQUIRE(npar >= 0 && npar <= kGMaxPar && (npar == 0 || (dops != nullptr && grad != nullptr)), FTLM_EINVAL,
                 "bad derivative operators");
    FTLM_REQUIRE(opts->start == FTLM_START_GAUSS_MT || opts->start == FTLM_START_PHASE, FTLM_EINVAL, "bad start");
    FTLM_REQUIRE(opts->reorth == FTLM_REORTH_NONE, FTLM_EUNSUPPORTED,
                 "derivatives follow the plain (unreorthogonalised) recurrence");
    for (int p = 0; p < npar; ++p)
      FTLM_REQUIRE(dops[p] != nullptr && dops[p]->dim == op->dim, FTLM_EINVAL, "derivative operator of another size");
    FTLM_HIP_CHECK(hipSetDevice(op->device));
    const int64_t dim = op->dim;
    const int R = std::max(1, opts->R);
    int M = std::max(1, opts->M);
    if (opts->clamp_M_to_dim) M = static_cast<int>(std::min<int64_t>(M, std::max<int64_t>(dim, 1)));
    FTLM_REQUIRE(R <= 32, FTLM_EUNSUPPORTED, "at most 32 samples per gradient call");
    FTLM_REQUIRE(M <= kGMaxM, FTLM_EUNSUPPORTED, "at most 512 Lanczos steps per gradient call");
    if (dim <= 0) {
      for (int b = 0; b < nb; ++b) lnZ[b] = -INFINITY;
      for (int i = 0; i < npar * nb; ++i) grad[i] = 0.0;
      return;
    }
    FTLM_REQUIRE(dim > 1, FTLM_EUNSUPPORTED, "a one-state operator has the closed form -beta H00");
    const hipStream_t st = nullptr;
    const int64_t n = dim * R;
    const size_t vb = static_cast<size_t>(n);
    // q, q', w and per parameter dq, dq', dw (= H dq, then the tangent of w) and tmp (= dH q)
    DevBuf<double2> q(vb), qp(vb), w(vb);
    std::vector<DevBuf<double2>> dq(npar), dqp(npar), dw(npar), tmp(npar);
    for (int p = 0; p < npar; ++p) {
      dq[p].alloc(vb);
      dqp[p].alloc(vb);
      dw[p].alloc(vb);
      tmp[p].alloc(vb);
      FTLM_HIP_CHECK(hipMemsetAsync(dq[p].p, 0, vb * sizeof(double2), st));
      FTLM_HIP_CHECK(hipMemsetAsync(dqp[p].p, 0, vb * sizeof(double2), st));
    }
    auto vecs = [&] {
      GradVecs V{};
      for (int p = 0; p < npar; ++p) {
        V.dq[p] = dq[p].p;
        V.dqp[p] = dqp[p].p;
        V.dw[p] = dw[p].p;
        V.tmp[p] = tmp[p].p;
      }
      return V;
    };
    // start block (src/lanczos.cpp:140-175), seeds seed + r * 100003
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
      std::vector<double2> h;
      gauss_interleaved(dim, R, seeds, &h);
      q.upload(h, st);
    }
    const int bsz = (kGBlock / R) * R;
    const int gblk = static_cast<int>(std::min<int64_t>(kGMaxGrid, (n + bsz - 1) / bsz));
    const int Q = 1 + npar;
    DevBuf<double2> part(static_cast<size_t>(gblk) * Q * R), dres(static_cast<size_t>(Q) * R);
    std::vector<double2> hres(static_cast<size_t>(Q) * R);
    // partials [block][quantity][R] -> hres[quantity * R + r]
    auto collect = [&](int nq) {
      hipLaunchKernelGGL(k_gsum, dim3(nq * R), dim3(kGBlock), 0, st, part.p, gblk, nq * R, dres.p);
      check_launch("k_gsum");
      FTLM_HIP_CHECK(hipMemcpyAsync(hres.data(), dres.p, sizeof(double2) * nq * R, hipMemcpyDeviceToHost, st));
      FTLM_HIP_CHECK(hipStreamSynchronize(st));
    };
    // normalise the start
    hipLaunchKernelGGL(k_gstep_dots, dim3(gblk), dim3(kGBlock), 0, st, n, R, 0, q.p, q.p, GradVecs{}, part.p);
    check_launch("k_gstep_dots");
    collect(1);
    for (int r = 0; r < R; ++r) FTLM_REQUIRE(std::sqrt(hres[r].x) >= 1e-18, FTLM_EUNSUPPORTED, "zero start vector");
    hipLaunchKernelGGL(k_gscale, dim3(gblk), dim3(kGBlock), 0, st, n, R, q.p, dres.p);
    check_launch("k_gscale");

    std::vector<double> ha(static_cast<size_t>(R) * M, 0.0), hb(static_cast<size_t>(R) * M, 0.0);
    std::vector<double> hda(static_cast<size_t>(npar) * R * M, 0.0), hdb(static_cast<size_t>(npar) * R * M, 0.0);
    std::vector<int> used(static_cast<size_t>(R), 0), alive(static_cast<size_t>(R), 1);
    std::vector<double> bprev(static_cast<size_t>(Q) * R, 0.0);  // b' and db'_p of the previous step
    std::vector<double2> z(static_cast<size_t>(Q) * R);
    std::vector<double> bet(static_cast<size_t>(Q) * R);
    DevBuf<double2> dz(static_cast<size_t>(Q) * R);
    DevBuf<double> dbp(static_cast<size_t>(Q) * R), dbt(static_cast<size_t>(Q) * R);
    DevBuf<int> dalive(static_cast<size_t>(R));
    for (int k = 0; k < M; ++k) {
      const GradVecs V = vecs();
      // w = H q;  dw_p = H dq_p;  tmp_p = dH_p q
      hv_apply(op, q.p, w.p, R, FTLM_C128, nullptr, nullptr, st);
      for (int p = 0; p < npar; ++p) {
        hv_apply(dops[p], q.p, tmp[p].p, R, FTLM_C128, nullptr, nullptr, st);
        hv_apply(op, dq[p].p, dw[p].p, R, FTLM_C128, nullptr, nullptr, st);
      }
      check_launch("gradient H*v");
      // z = <q|w>;  dz_p = <dq_p|w> + <q|dw_p + tmp_p>
      hipLaunchKernelGGL(k_gstep_dots, dim3(gblk), dim3(kGBlock), 0, st, n, R, npar, q.p, w.p, V, part.p);
      check_launch("k_gstep_dots");
      collect(Q);
      for (int r = 0; r < R; ++r)
        for (int j = 0; j < Q; ++j) {
          const size_t o = static_cast<size_t>(j) * R + r;
          z[o] = alive[r] ? hres[o] : make_double2(0.0, 0.0);
          if (!alive[r]) continue;
          if (j == 0)
            ha[static_cast<size_t>(r) * M + k] = hres[o].x;
          else
            hda[(static_cast<size_t>(j - 1) * R + r) * M + k] = hres[o].x;
        }
      dz.upload(z, st);
      dbp.upload(bprev, st);
      dalive.upload(alive, st);
      hipLaunchKernelGGL(k_gstep_update, dim3(gblk), dim3(kGBlock), 0, st, n, R, npar, q.p, qp.p, w.p, V, dz.p, dbp.p,
                         dalive.p, k > 0 ? 1 : 0, part.p);
      check_launch("k_gstep_update");
      collect(Q);  // |w|^2 and Re <w|dw_p>
      std::fill(bet.begin(), bet.end(), 0.0);
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
          const size_t o = static_cast<size_t>(1 + p) * R + r;
          const double db = hres[o].x / b;
          hdb[(static_cast<size_t>(p) * R + r) * M + k] = db;
          bet[o] = db;
          bprev[o] = db;
        }
      }
      bool any = false;
      for (int r = 0; r < R; ++r) any = any || alive[r];
      if (!any || k + 1 == M) break;
      dbt.upload(bet, st);
      dalive.upload(alive, st);
      hipLaunchKernelGGL(k_gstep_norm, dim3(gblk), dim3(kGBlock), 0, st, n, R, npar, w.p, V, dbt.p, dalive.p);
      check_launch("k_gstep_norm");
      // rotate: (q', q, w) <- (q, w, q') and likewise for the tangents
      std::swap(qp.p, q.p);
      std::swap(q.p, w.p);
      for (int p = 0; p < npar; ++p) {
        std::swap(dqp[p].p, dq[p].p);
        std::swap(dq[p].p, dw[p].p);
      }
    }
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
    // quadrature and its derivatives, one workgroup per sample
    DevBuf<double> da(ha.size()), db(hb.size()), dda(std::max<size_t>(hda.size(), 1)), ddb(std::max<size_t>(hdb.size(), 1));
    da.upload(ha, st);
    db.upload(hb, st);
    if (npar) {
      dda.upload(hda, st);
      ddb.upload(hdb, st);
    }
    DevBuf<int> dn(static_cast<size_t>(R));
    dn.upload(used, st);
    DevBuf<double> dbeta(static_cast<size_t>(nb));
    dbeta.upload(betas, static_cast<size_t>(nb), st);
    DevBuf<double> Uw(static_cast<size_t>(R) * M * M), Wm(static_cast<size_t>(R) * M * M);
    DevBuf<double> Ls(static_cast<size_t>(R) * nb), dLs(static_cast<size_t>(std::max(npar, 1)) * R * nb);
    DevBuf<int> status(1);
    FTLM_HIP_CHECK(hipMemsetAsync(status.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_grad_tridiag, dim3(R), dim3(kGBlock), 0, st, da.p, db.p, dda.p, ddb.p, dn.p, R, M, npar,
                       dbeta.p, nb, Uw.p, Wm.p, Ls.p, dLs.p, static_cast<double*>(nullptr), status.p);
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


int ftlm_hip_ftlm_tape_forward(const ftlm_op_t* op, ftlm_op_t* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                               const double* betas, int nb, int ckpt, int mode, double* lnZ, ftlm_tape_t** out) {
  return guarded([&] {
    FTLM_REQUIRE(out != nullptr, FTLM_EINVAL, "null output handle");
    *out = nullptr;
    auto T = std::make_unique<ftlm_tape>();
    tape_forward(T.get(), op, dops, npar, opts, betas, nb, ckpt, mode, lnZ);
    *out = T.release();
  });
}

int ftlm_hip_ftlm_tape_backward(const ftlm_tape_t* tape, const double* gw, double* grad) {
  return guarded([&] {
    FTLM_REQUIRE(tape != nullptr, FTLM_EINVAL, "null tape");
    tape_backward(tape, gw, grad);
  });
}

int ftlm_hip_ftlm_tape_destroy(ftlm_tape_t* tape) {
  return guarded([&] { delete tape; });
}

int ftlm_hip_ftlm_observables(const ftlm_op_t* op, ftlm_op_t* const* aops, int nobs, const ftlm_ftlm_opts_t* opts,
                              const double* betas, int nb, double* lnZ, double* obs) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && aops != nullptr && obs != nullptr && lnZ != nullptr && betas != nullptr && nb >= 1,
                 FTLM_EINVAL, "null argument");
    FTLM_REQUIRE(nobs >= 1 && nobs <= kGMaxPar, FTLM_EINVAL, "between 1 and 16 observables");
    for (int p = 0; p < nobs; ++p)
      FTLM_REQUIRE(aops[p] != nullptr && aops[p]->dim == op->dim, FTLM_EINVAL, "observable of another size");
    ftlm_tape T;
    ObsAcc acc;
    acc.aops.assign(aops, aops + nobs);
    tape_forward(&T, op, nullptr, 0, opts, betas, nb, std::max(1, opts->M), FTLM_TS_GENERIC, lnZ, &acc);
    const int R = T.R, M = T.M;
    const hipStream_t st = nullptr;
    std::vector<double> ha(static_cast<size_t>(R) * M, 0.0), hb(static_cast<size_t>(R) * M, 0.0);
    std::vector<double2> hA(static_cast<size_t>(nobs) * R * M, make_double2(0.0, 0.0));
    for (int r = 0; r < R; ++r)
      for (int k = 0; k < T.used[r]; ++k) {
        ha[static_cast<size_t>(r) * M + k] = T.zs[static_cast<size_t>(k) * R + r].x;
        hb[static_cast<size_t>(r) * M + k] = T.bs[static_cast<size_t>(k) * R + r];
        for (int p = 0; p < nobs; ++p)
          hA[(static_cast<size_t>(p) * R + r) * M + k] = acc.a[(static_cast<size_t>(p) * M + k) * R + r];
      }
    DevBuf<double> da(ha.size()), db(hb.size()), dbeta(static_cast<size_t>(nb)), Uw(static_cast<size_t>(R) * M * M);
    DevBuf<double> Ls(static_cast<size_t>(R) * nb), num(static_cast<size_t>(nobs) * R * nb);
    DevBuf<double2> dA(hA.size());
    DevBuf<int> dn(static_cast<size_t>(R)), status(1);
    da.upload(ha, st);
    db.upload(hb, st);
    dA.upload(hA, st);
    dbeta.upload(betas, static_cast<size_t>(nb), st);
    dn.upload(T.used, st);
    FTLM_HIP_CHECK(hipMemsetAsync(status.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_obs_tridiag, dim3(R), dim3(kGBlock), 0, st, da.p, db.p, dn.p, R, M, nobs, dA.p, dbeta.p, nb,
                       Uw.p, Ls.p, num.p, status.p);
    check_launch("k_obs_tridiag");
    int bad = 0;
    FTLM_HIP_CHECK(hipMemcpy(&bad, status.p, sizeof(int), hipMemcpyDeviceToHost));
    FTLM_REQUIRE(bad == 0, FTLM_EHIP, "tridiagonal QL did not converge");
    std::vector<double> hL(Ls.n), hN(num.n);
    FTLM_HIP_CHECK(hipMemcpy(hL.data(), Ls.p, hL.size() * sizeof(double), hipMemcpyDeviceToHost));
    FTLM_HIP_CHECK(hipMemcpy(hN.data(), num.p, hN.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int b = 0; b < nb; ++b) {  // <A>(beta) = sum_r softmax_r(L_r(beta)) num_r
      double mx = -INFINITY;
      for (int r = 0; r < R; ++r) mx = std::max(mx, hL[static_cast<size_t>(r) * nb + b]);
      double zsum = 0.0;
      for (int r = 0; r < R; ++r) zsum += std::exp(hL[static_cast<size_t>(r) * nb + b] - mx);
      for (int p = 0; p < nobs; ++p) {
        double v = 0.0;
        for (int r = 0; r < R; ++r)
          v += std::exp(hL[static_cast<size_t>(r) * nb + b] - mx) / zsum *
               hN[(static_cast<size_t>(p) * R + r) * nb + b];
        obs[static_cast<size_t>(p) * nb + b] = v;
      }
    }
  });
}

int ftlm_hip_ftlm_vjp(const ftlm_op_t* op, ftlm_op_t* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                      const double* betas, int nb, const double* gw, int ckpt, int mode, double* lnZ, double* grad) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && gw != nullptr && lnZ != nullptr && (npar == 0 || grad != nullptr), FTLM_EINVAL,
                 "null argument");
    if (op->dim <= 0) {
      for (int b = 0; b < nb; ++b) lnZ[b] = -INFINITY;
      for (int i = 0; i < npar; ++i) grad[i] = 0.0;
      return;
    }
    ftlm_tape T;
    tape_forward(&T, op, dops, npar, opts, betas, nb, ckpt, mode, lnZ);
    tape_backward(&T, gw, grad);
  });
}

}  // extern "C"
