// tridiag.hip -- Gauss-quadrature eigensolvers for Lanczos tridiagonals and the FTLM Ritz-weight
// thermodynamic sums, on the device.
//
// k_jacobi: the reference's max-pivot single-rotation Jacobi (src/ftlm_thermo.cpp:17-80 with
//   max(3000,120n) rotations at 1e-14; src/lanczos.cpp:38-102 with 120 rotations at 1e-15 for
//   the extrema).  One workgroup per tridiagonal.  The pivot search keeps, per row i, the largest
//   |A[i][j]| (j > i) and its first column; a rotation on (p, q) changes rows p, q entirely and
//   other rows only in columns p and q, so each row is rescanned only when its maximum may have
//   dropped.  The global pivot is the first strict maximum in row-major order, exactly as the
//   reference's nested scan.  Only row 0 of the eigenvector matrix is accumulated: it is all the
//   quadrature weights need, and each row of V is rotated independently, so the values equal the
//   reference's V[0][k].
// k_ql: implicit-shift QL (Wilkinson shift) run to convergence, one thread per tridiagonal;
//   the converged alternative (FTLM_QUAD_CONVERGED).
// k_thermo_op / k_thermo_gc: log-sum-exp Ritz sums for lnZ(beta) and E(beta) per operator and
//   the grand-canonical (beta, mu) mixture (src/ftlm_thermo.cpp:120-145, 195-210;
//   benchmarks/bench_ftlm_nmu_rect.cpp:186-219).

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <cmath>

#include "common.hpp"

namespace ftlm_hip {

constexpr int kJacBlock = 256;

struct RowMax {
  double v;
  int col;
};

// Rows of the pivot table to rescan are handled one per wavefront: the lanes stride over the row
// and combine (larger |a|, then smaller column), which is the serial scan's first strict maximum.
__device__ __forceinline__ void rescan_rows(const double* A, int n, RowMax* rm, const int* todo, int ntodo) {
  const int lane = static_cast<int>(threadIdx.x) & 63, wv = static_cast<int>(threadIdx.x) >> 6;
  const int nw = static_cast<int>(blockDim.x) >> 6;
  for (int t = wv; t < ntodo; t += nw) {
    const int i = todo[t];
    double bv = 0.0;
    int bc = -1;
    for (int j = i + 1 + lane; j < n; j += 64) {
      const double v = fabs(A[static_cast<int64_t>(i) * n + j]);
      if (v > bv) {
        bv = v;
        bc = j;
      }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int oc = __shfl_xor(bc, o, 64);
      if (ov > bv || (ov == bv && oc >= 0 && (bc < 0 || oc < bc))) {
        bv = ov;
        bc = oc;
      }
    }
    if (lane == 0) rm[i] = RowMax{bv, bc};
  }
}

// Dynamic LDS of k_jacobi for matrices of at most n rows: v0 [n], row maxima [n], rescan list [n],
// and the matrix itself [n][n] when it is kept in LDS.
constexpr int kJacLdsMaxN = 85;  // n * n * 8 + 28 n <= 60 KB
__host__ __device__ inline size_t jacobi_lds_offset(int n) {
  return (static_cast<size_t>(n) * (sizeof(double) + sizeof(RowMax) + sizeof(int)) + 15) / 16 * 16;
}
inline size_t jacobi_shmem(int n, bool lds) {
  return jacobi_lds_offset(n) + (lds ? static_cast<size_t>(n) * n * sizeof(double) : 0);
}

// One tridiagonal per workgroup.  alpha/beta are strided: entry i of sample s at [i*stride + s].
// n[s] = size (used steps), nb_avail = number of valid beta entries (n-1 or fewer).  kLds keeps
// the matrix in LDS (n <= kJacLdsMaxN), otherwise in the workgroup's slice of `scratch`.  Every
// rotation runs the same arithmetic and pivot order either way.
template <bool kLds>
__global__ __launch_bounds__(kJacBlock) void k_jacobi(const double* __restrict__ alpha, const double* __restrict__ beta,
                                                      const int* __restrict__ nvec, const int* __restrict__ nbvec,
                                                      int stride, int cap_fixed, double tol, int with_weights,
                                                      double* __restrict__ scratch, int64_t scratch_per,
                                                      double* __restrict__ ritz, double* __restrict__ weight,
                                                      int out_stride, const double* __restrict__ dense) {
#pragma clang fp contract(off)
  const int s = blockIdx.x;
  const int n = nvec[s];
  const int nb = nbvec ? nbvec[s] : (n - 1);
  extern __shared__ unsigned char smem_raw[];
  double* v0 = reinterpret_cast<double*>(smem_raw);                 // [n]
  RowMax* rm = reinterpret_cast<RowMax*>(v0 + n);                    // [n]
  int* todo = reinterpret_cast<int*>(rm + n);                        // [n]
  double* A = kLds ? reinterpret_cast<double*>(smem_raw + jacobi_lds_offset(n)) : scratch + scratch_per * s;
  __shared__ double s_val[kJacBlock / 64];
  __shared__ int s_row[kJacBlock / 64];
  __shared__ double s_c, s_s, s_big;
  __shared__ int s_p, s_q, s_ntodo;
  const int tid = threadIdx.x;
  if (n <= 0) return;

  if (dense) {  // a full symmetric n x n matrix per workgroup (row-major, n = nvec[s])
    for (int64_t k = tid; k < static_cast<int64_t>(n) * n; k += blockDim.x) A[k] = dense[scratch_per * s + k];
    for (int i = tid; i < n; i += blockDim.x) v0[i] = (i == 0) ? 1.0 : 0.0;
  } else {
    for (int64_t k = tid; k < static_cast<int64_t>(n) * n; k += blockDim.x) A[k] = 0.0;
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
      A[static_cast<int64_t>(i) * n + i] = alpha[static_cast<int64_t>(i) * stride + s];
      if (i < n - 1 && i < nb) {
        const double b = beta[static_cast<int64_t>(i) * stride + s];
        A[static_cast<int64_t>(i) * n + i + 1] = b;
        A[static_cast<int64_t>(i + 1) * n + i] = b;
      }
      v0[i] = (i == 0) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  const int cap = cap_fixed > 0 ? cap_fixed : max(3000, 120 * n);
  if (n > 1) {
    for (int i = tid; i < n; i += blockDim.x) todo[i] = i;
    __syncthreads();
    rescan_rows(A, n, rm, todo, n);
    __syncthreads();
    for (int it = 0; it < cap; ++it) {
      // ---- pivot: largest row maximum, smallest row on ties
      double bv = -1.0;
      int br = n;
      for (int i = tid; i < n; i += blockDim.x) {
        const double v = rm[i].col >= 0 ? rm[i].v : 0.0;
        if (v > bv || (v == bv && i < br)) {
          bv = v;
          br = i;
        }
      }
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int orow = __shfl_xor(br, o, 64);
        if (ov > bv || (ov == bv && orow < br)) {
          bv = ov;
          br = orow;
        }
      }
      if ((tid & 63) == 0) {
        s_val[tid >> 6] = bv;
        s_row[tid >> 6] = br;
      }
      __syncthreads();
      if (tid == 0) {
        double v = s_val[0];
        int r = s_row[0];
        for (int w = 1; w < kJacBlock / 64; ++w)
          if (s_val[w] > v || (s_val[w] == v && s_row[w] < r)) {
            v = s_val[w];
            r = s_row[w];
          }
        // the reference's scan starts from (p, q) = (0, 1), max 0: a zero maximum keeps (0, 1)
        int p = 0, q = 1;
        if (v > 0.0 && r < n && rm[r].col >= 0) {
          p = r;
          q = rm[r].col;
        } else {
          v = 0.0;
        }
        s_big = v;
        s_p = p;
        s_q = q;
        s_ntodo = 0;
        if (v >= tol) {
          const double app = A[static_cast<int64_t>(p) * n + p];
          const double aqq = A[static_cast<int64_t>(q) * n + q];
          const double apq = A[static_cast<int64_t>(p) * n + q];
          const double phi = 0.5 * atan2(2.0 * apq, aqq - app);
          const double c = cos(phi), sn = sin(phi);
          s_c = c;
          s_s = sn;
          const double npp = c * c * app - 2.0 * sn * c * apq + sn * sn * aqq;
          const double nqq = sn * sn * app + 2.0 * sn * c * apq + c * c * aqq;
          // diagonal / pivot entries are written after the off-diagonal sweep below reads rows
          // p and q, so stash them
          A[static_cast<int64_t>(p) * n + q] = npp;  // temporarily hold npp in (p,q)
          A[static_cast<int64_t>(q) * n + p] = nqq;  // and nqq in (q,p)
          const double vp = v0[p], vq = v0[q];
          v0[p] = c * vp - sn * vq;
          v0[q] = sn * vp + c * vq;
        }
      }
      __syncthreads();
      if (s_big < tol) break;
      const int p = s_p, q = s_q;
      const double c = s_c, sn = s_s;
      for (int k = tid; k < n; k += blockDim.x) {
        if (k == p || k == q) continue;
        const double apk = A[static_cast<int64_t>(p) * n + k];
        const double aqk = A[static_cast<int64_t>(q) * n + k];
        const double rp = c * apk - sn * aqk;
        const double rq = c * aqk + sn * apk;
        A[static_cast<int64_t>(p) * n + k] = rp;
        A[static_cast<int64_t>(k) * n + p] = rp;
        A[static_cast<int64_t>(q) * n + k] = rq;
        A[static_cast<int64_t>(k) * n + q] = rq;
      }
      __syncthreads();
      if (tid == 0) {
        const double npp = A[static_cast<int64_t>(p) * n + q];
        const double nqq = A[static_cast<int64_t>(q) * n + p];
        A[static_cast<int64_t>(p) * n + q] = 0.0;
        A[static_cast<int64_t>(q) * n + p] = 0.0;
        A[static_cast<int64_t>(p) * n + p] = npp;
        A[static_cast<int64_t>(q) * n + q] = nqq;
      }
      __syncthreads();
      // ---- row maxima: rows p, q fully; rows i < q see new values in columns p (if i < p), q;
      // rows whose maximum sat in column p or q are rescanned (by wavefronts, below)
      for (int i = tid; i < n; i += blockDim.x) {
        if (i == p || i == q) {
          todo[atomicAdd(&s_ntodo, 1)] = i;
          continue;
        }
        if (i > q && i > p) continue;
        RowMax r = rm[i];
        if (r.col == p || r.col == q) {
          todo[atomicAdd(&s_ntodo, 1)] = i;
          continue;
        }
        const int c1 = p < q ? p : q, c2 = p < q ? q : p;
        if (c1 > i) {
          const double v = fabs(A[static_cast<int64_t>(i) * n + c1]);
          if (v > r.v || (v == r.v && r.col >= 0 && c1 < r.col) || (r.col < 0 && v > 0.0)) {
            r.v = v;
            r.col = c1;
          }
        }
        if (c2 > i) {
          const double v = fabs(A[static_cast<int64_t>(i) * n + c2]);
          if (v > r.v || (v == r.v && r.col >= 0 && c2 < r.col) || (r.col < 0 && v > 0.0)) {
            r.v = v;
            r.col = c2;
          }
        }
        rm[i] = r;
      }
      __syncthreads();
      rescan_rows(A, n, rm, todo, s_ntodo);
      __syncthreads();
    }
  }
  for (int k = tid; k < n; k += blockDim.x) {
    ritz[static_cast<int64_t>(s) * out_stride + k] = A[static_cast<int64_t>(k) * n + k];
    if (with_weights) weight[static_cast<int64_t>(s) * out_stride + k] = v0[k] * v0[k];
  }
}

// k_jacobi over `grid` problems of at most max_n rows, in LDS when they fit.
void launch_jacobi(int grid, int max_n, hipStream_t st, const double* alpha, const double* beta, const int* nvec,
                   const int* nbvec, int stride, int cap_fixed, double tol, int with_weights, double* scratch,
                   int64_t scratch_per, double* ritz, double* weight, int out_stride, const double* dense) {
  const bool lds = max_n <= kJacLdsMaxN;
  const size_t shmem = jacobi_shmem(max_n, lds);
  FTLM_REQUIRE(shmem <= 60 * 1024, FTLM_EUNSUPPORTED, "matrix too large for the Jacobi kernel");
  if (lds)
    hipLaunchKernelGGL(k_jacobi<true>, dim3(grid), dim3(kJacBlock), shmem, st, alpha, beta, nvec, nbvec, stride,
                       cap_fixed, tol, with_weights, scratch, scratch_per, ritz, weight, out_stride, dense);
  else
    hipLaunchKernelGGL(k_jacobi<false>, dim3(grid), dim3(kJacBlock), shmem, st, alpha, beta, nvec, nbvec, stride,
                       cap_fixed, tol, with_weights, scratch, scratch_per, ritz, weight, out_stride, dense);
  check_launch("k_jacobi");
}

// Converged eigen-decomposition of a symmetric tridiagonal (implicit QL, Wilkinson shift),
// returning eigenvalues and squared first components of the eigenvectors.
__global__ void k_ql(const double* __restrict__ alpha, const double* __restrict__ beta, const int* __restrict__ nvec,
                     const int* __restrict__ nbvec, int stride, int nsamp, int max_n, double* __restrict__ ritz,
                     double* __restrict__ weight, int out_stride, int* __restrict__ status) {
  // d, e and the first eigenvector row z of every thread's tridiagonal live in LDS, interleaved
  // by thread ([i * blockDim + tid]), so the sequential QL sweeps never wait on global memory
  extern __shared__ double ql_lds[];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int nt = blockDim.x, t = threadIdx.x;
  double* dl = ql_lds;
  double* el = ql_lds + static_cast<int64_t>(max_n) * nt;
  double* zl = ql_lds + 2 * static_cast<int64_t>(max_n) * nt;
#define QD(i) dl[(i) * nt + t]
#define QE(i) el[(i) * nt + t]
#define QZ(i) zl[(i) * nt + t]
  if (s >= nsamp) return;
  const int n = nvec[s];
  const int nb = nbvec ? nbvec[s] : n - 1;
  for (int i = 0; i < n; ++i) {
    QD(i) = alpha[static_cast<int64_t>(i) * stride + s];
    QE(i) = (i < n - 1 && i < nb) ? beta[static_cast<int64_t>(i) * stride + s] : 0.0;
    QZ(i) = (i == 0) ? 1.0 : 0.0;
  }
  int bad = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = fabs(QD(m)) + fabs(QD(m + 1));
        if (fabs(QE(m)) <= 1e-300 + 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) {
        bad = 1;
        break;
      }
      double g = (QD(l + 1) - QD(l)) / (2.0 * QE(l));
      double r = hypot(g, 1.0);
      g = QD(m) - QD(l) + QE(l) / (g + (g >= 0.0 ? r : -r));
      double sn = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      // d[i+1] and z[i+1] are carried in registers (no rotation of this sweep has changed d[i+1]
      // since it was read as d[i]; z[i+1] is the previous rotation's z[i]); e, d, z at i-1 are
      // loaded while rotation i runs.  Entries are bounded (tridiagonals of Lanczos on a
      // Hamiltonian), so r = sqrt(f^2 + g^2) needs no scaling.
      double d_n1 = QD(m), z_n1 = QZ(m);
      double e_i = QE(m - 1), d_i = QD(m - 1), z_i = QZ(m - 1);
      for (int i = m - 1; i >= l; --i) {
        double e_p = 0.0, d_p = 0.0, z_p = 0.0;
        if (i > l) {
          e_p = QE(i - 1);
          d_p = QD(i - 1);
          z_p = QZ(i - 1);
        }
        const double f = sn * e_i;
        const double b = c * e_i;
        r = sqrt(f * f + g * g);
        QE(i + 1) = r;
        if (r == 0.0) {
          QD(i + 1) = d_n1 - p;
          QE(m) = 0.0;
          deflated = true;
          break;
        }
        const double inv = 1.0 / r;
        sn = f * inv;
        c = g * inv;
        g = d_n1 - p;
        r = (d_i - g) * sn + 2.0 * c * b;
        p = sn * r;
        QD(i + 1) = g + p;
        g = c * r - b;
        QZ(i + 1) = sn * z_i + c * z_n1;
        const double zi = c * z_i - sn * z_n1;
        QZ(i) = zi;
        d_n1 = d_i;
        z_n1 = zi;
        e_i = e_p;
        d_i = d_p;
        z_i = z_p;
      }
      if (deflated) continue;
      QD(l) -= p;
      QE(l) = g;
      QE(m) = 0.0;
    }
  }
  double* d = ritz + static_cast<int64_t>(s) * out_stride;
  double* z = weight + static_cast<int64_t>(s) * out_stride;
  for (int i = 0; i < n; ++i) {
    d[i] = QD(i);
    z[i] = QZ(i) * QZ(i);
  }
#undef QD
#undef QE
#undef QZ
  if (bad) atomicAdd(status, 1);
}

// ln sum_k w_k exp(-beta lambda_k) over k < n, skipping w <= 0 (src/ftlm_thermo.cpp:120-145).
__device__ double log_quad(const double* lam, const double* w, int n, double bt, double* e_num) {
  if (n <= 0) return -INFINITY;
  if (n == 1) {
    if (e_num) *e_num = lam[0];
    return -bt * lam[0];
  }
  double mx = -INFINITY;
  for (int k = 0; k < n; ++k)
    if (w[k] > 0.0) mx = fmax(mx, -bt * lam[k] + log(w[k]));
  if (!isfinite(mx)) return -INFINITY;
  double s = 0.0, se = 0.0;
  for (int k = 0; k < n; ++k) {
    if (w[k] <= 0.0) continue;
    const double t = exp(-bt * lam[k] + log(w[k]) - mx);
    s += t;
    se += t * lam[k];
  }
  if (e_num) *e_num = se / s;  // energy of this sample's quadrature
  return mx + log(s);
}

// Per operator and beta: lnZ = ln dim + log-mean-exp_r L_r; E = sum_r e^{L_r} E_r / sum_r e^{L_r}.
__global__ void k_thermo_op(const double* ritz, const double* weight, const int* used, int n_ops, int R, int M,
                            const int64_t* dims, const double* betas, int nb, double* lnZ, double* E) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_ops * nb) return;
  const int op = t / nb, ib = t % nb;
  const double bt = betas[ib];
  double mx = -INFINITY;
  int cnt = 0;
  for (int r = 0; r < R; ++r) {
    const int n = used[op * R + r];
    if (n <= 0) continue;
    const int64_t o = (static_cast<int64_t>(op) * R + r) * M;
    const double L = log_quad(ritz + o, weight + o, n, bt, nullptr);
    if (cnt == 0 || L > mx) mx = L;
    ++cnt;
  }
  double lz = -INFINITY, en = NAN;
  if (cnt > 0 && dims[op] > 0) {
    double s = 0.0, se = 0.0;
    for (int r = 0; r < R; ++r) {
      const int n = used[op * R + r];
      if (n <= 0) continue;
      const int64_t o = (static_cast<int64_t>(op) * R + r) * M;
      double er = 0.0;
      const double L = log_quad(ritz + o, weight + o, n, bt, &er);
      const double wgt = exp(L - mx);
      s += wgt;
      se += wgt * er;
    }
    lz = log(static_cast<double>(dims[op])) + (mx + log(s / static_cast<double>(cnt)));
    en = se / s;
  }
  if (lnZ) lnZ[t] = lz;
  if (E) E[t] = en;
}

// Grand canonical mixture over operators (sectors): weight_s = exp(lnZ_s + beta mu N_s).
__global__ void k_thermo_gc(const double* lnZ, const double* E, const int* N, int n_ops, const double* betas, int nb,
                            const double* mus, int nmu, double* lnZ_gc, double* N_gc, double* E_gc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * nmu) return;
  const int ib = t / nmu, im = t % nmu;
  const double bt = betas[ib], mu = mus[im];
  double mx = -INFINITY;
  for (int s = 0; s < n_ops; ++s) {
    const double l = lnZ[s * nb + ib];
    if (!isfinite(l)) continue;
    mx = fmax(mx, l + bt * mu * static_cast<double>(N[s]));
  }
  double z = 0.0, zn = 0.0, ze = 0.0;
  if (isfinite(mx)) {
    for (int s = 0; s < n_ops; ++s) {
      const double l = lnZ[s * nb + ib];
      if (!isfinite(l)) continue;
      const double w = exp(l + bt * mu * static_cast<double>(N[s]) - mx);
      z += w;
      zn += static_cast<double>(N[s]) * w;
      if (E) ze += E[s * nb + ib] * w;
    }
  }
  if (lnZ_gc) lnZ_gc[t] = isfinite(mx) ? mx + log(z) : -INFINITY;
  if (N_gc) N_gc[t] = z > 0.0 ? zn / z : 0.0;
  if (E_gc) E_gc[t] = z > 0.0 ? ze / z : NAN;
}

// Real 2d x 2d embedding of a complex d x d block, [[Re, -Im], [Im, Re]]
// (benchmarks/bench_ftlm_nmu_4x4_k.cpp:298-311); H is row-major complex.
__global__ void k_embed_real(const double2* __restrict__ H, int d, double* __restrict__ R) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(d) * d) return;
  const int i = static_cast<int>(t / d), j = static_cast<int>(t % d);
  const double2 h = H[t];
  const int64_t m = 2 * d;
  R[i * m + j] = h.x;
  R[i * m + j + d] = -h.y;
  R[(i + d) * m + j] = h.y;
  R[(i + d) * m + j + d] = h.x;
}

__global__ void k_identity_block(double2* X, int d) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(d) * d) return;
  X[t] = make_double2((t / d) == (t % d) ? 1.0 : 0.0, 0.0);
}

// ln Z = ln(1/2) + max_t + ln sum_i exp(t_i - max_t), t_i = -beta lambda_i over the 2d
// eigenvalues of the embedding (every eigenvalue of H twice).
__global__ void k_ed_lnz(const double* lam, int m, const double* betas, int nb, double* lnZ) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double bt = betas[b];
  double mx = -INFINITY;
  for (int i = 0; i < m; ++i) mx = fmax(mx, -bt * lam[i]);
  double sum = 0.0;
  for (int i = 0; i < m; ++i) sum += exp(-bt * lam[i] - mx);
  lnZ[b] = log(0.5) + mx + log(sum);
}

// ---------------------------------------------------------------- host drivers

// Diagonalise nsamp tridiagonals stored strided in device memory.  Outputs (device) ritz and
// weight [nsamp][out_stride].
void quadrature_device(const double* d_alpha, const double* d_beta, const int* d_n, const int* d_nb, int stride,
                       int nsamp, int max_n, int quad, int cap_fixed, double tol, int with_weights, double* d_ritz,
                       double* d_weight, int out_stride, hipStream_t st) {
  if (nsamp <= 0 || max_n <= 0) return;
  if (quad == FTLM_QUAD_CONVERGED) {
    DevBuf<int> status(1);
    FTLM_HIP_CHECK(hipMemsetAsync(status.p, 0, sizeof(int), st));
    // threads per workgroup: as many tridiagonals as fit 3 * max_n doubles each in the LDS budget
    int dev = 0, lds_max = 0;
    FTLM_HIP_CHECK(hipGetDevice(&dev));
    FTLM_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t budget = std::min<size_t>(120 * 1024, static_cast<size_t>(lds_max));
    int nt = 64;
    while (nt > 1 && static_cast<size_t>(nt) * 3 * max_n * sizeof(double) > budget) nt /= 2;
    const size_t shmem = static_cast<size_t>(nt) * 3 * max_n * sizeof(double);
    FTLM_REQUIRE(shmem <= budget, FTLM_EUNSUPPORTED, "tridiagonal too large for the QL kernel");
    hipLaunchKernelGGL(k_ql, dim3(grid_for(nsamp, nt)), dim3(nt), shmem, st, d_alpha, d_beta, d_n, d_nb, stride, nsamp,
                       max_n, d_ritz, d_weight, out_stride, status.p);
    check_launch("k_ql");
    int bad = 0;
    FTLM_HIP_CHECK(hipMemcpyAsync(&bad, status.p, sizeof(int), hipMemcpyDeviceToHost, st));
    FTLM_HIP_CHECK(hipStreamSynchronize(st));
    FTLM_REQUIRE(bad == 0, FTLM_EHIP, "tridiagonal QL did not converge");
    return;
  }
  const int64_t per = static_cast<int64_t>(max_n) * max_n;
  DevBuf<double> scratch(static_cast<size_t>(per) * nsamp);
  launch_jacobi(nsamp, max_n, st, d_alpha, d_beta, d_n, d_nb, stride, cap_fixed, tol, with_weights, scratch.p, per,
                d_ritz, d_weight, out_stride, nullptr);
  FTLM_HIP_CHECK(hipStreamSynchronize(st));
}

void thermo_device(const double* d_ritz, const double* d_weight, const int* d_used, int n_ops, int R, int M,
                   const int64_t* d_dims, const int* d_N, const double* d_betas, int nb, const double* d_mus, int nmu,
                   double* d_lnZ, double* d_E, double* d_lnZgc, double* d_Ngc, double* d_Egc, hipStream_t st) {
  hipLaunchKernelGGL(k_thermo_op, dim3(grid_for(static_cast<int64_t>(n_ops) * nb, 64)), dim3(64), 0, st, d_ritz,
                     d_weight, d_used, n_ops, R, M, d_dims, d_betas, nb, d_lnZ, d_E);
  check_launch("k_thermo_op");
  if (nmu > 0 && d_N != nullptr) {
    hipLaunchKernelGGL(k_thermo_gc, dim3(grid_for(static_cast<int64_t>(nb) * nmu, 64)), dim3(64), 0, st, d_lnZ, d_E,
                       d_N, n_ops, d_betas, nb, d_mus, nmu, d_lnZgc, d_Ngc, d_Egc);
    check_launch("k_thermo_gc");
  }
}

// ------------------------------------------------------------------ allocation-free thermodynamics
// Device workspace of ftlm_hip_ftlm_thermo_noalloc: one file-scope array per device (the loader
// gives every device its own copy; it lives until the process ends).  Host code never uses the
// array's name as an address: the per-device address comes from hipGetSymbolAddress.
constexpr size_t kThermoScratchBytes = size_t{16} << 20;
static __device__ __align__(256) unsigned char scratch[kThermoScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;  // device -> address of that device's `scratch`

// Address of the current device's workspace, looked up once per device.
unsigned char* thermo_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    throw Error(FTLM_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  return static_cast<unsigned char*>(p);
}

// Bump allocation of 256-byte aligned pieces of the workspace (no device allocation).
struct ScratchCarve {
  unsigned char* base;
  size_t used = 0;
  template <class T>
  T* take(size_t count) {
    const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
    FTLM_REQUIRE(used + bytes <= kThermoScratchBytes, FTLM_EUNSUPPORTED,
                 "ftlm_hip_ftlm_thermo_noalloc: inputs exceed the " + std::to_string(kThermoScratchBytes >> 20) +
                     " MB device workspace (use ftlm_hip_ftlm_thermo)");
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};

}  // namespace

}  // namespace ftlm_hip

using namespace ftlm_hip;

extern "C" {

int ftlm_hip_exact_log_partition(const ftlm_op_t* op, const double* betas, int nb, double* lnZ, double* evals) {
  return guarded([&] {
    FTLM_REQUIRE(op != nullptr && betas != nullptr && lnZ != nullptr && nb >= 1, FTLM_EINVAL, "bad arguments");
    FTLM_HIP_CHECK(hipSetDevice(op->device));
    const int64_t d64 = op->dim;
    if (d64 <= 0) {
      for (int b = 0; b < nb; ++b) lnZ[b] = -INFINITY;
      return;
    }
    FTLM_REQUIRE(d64 <= 1024, FTLM_EUNSUPPORTED, "exact diagonalisation is for blocks of at most 1024 states");
    const int d = static_cast<int>(d64);
    // dense H = H * I: the block apply of the d unit vectors, row-major in the vector-block layout
    DevBuf<double2> X(static_cast<size_t>(d) * d), Y(static_cast<size_t>(d) * d);
    hipLaunchKernelGGL(k_identity_block, dim3(grid_for(static_cast<int64_t>(d) * d, 256)), dim3(256), 0, nullptr, X.p, d);
    check_launch("k_identity_block");
    hv_apply(op, X.p, Y.p, d, FTLM_C128, nullptr, nullptr, nullptr);
    if (d == 1) {  // a single state: ln Z = -beta H00 (bench_ftlm_nmu_4x4_k.cpp:254-260)
      double2 h;
      FTLM_HIP_CHECK(hipMemcpy(&h, Y.p, sizeof(double2), hipMemcpyDeviceToHost));
      for (int b = 0; b < nb; ++b) lnZ[b] = -betas[b] * h.x;
      if (evals) evals[0] = evals[1] = h.x;
      return;
    }
    const int m = 2 * d;
    DevBuf<double> Rm(static_cast<size_t>(m) * m), scratch(static_cast<size_t>(m) * m), lam(static_cast<size_t>(m));
    hipLaunchKernelGGL(k_embed_real, dim3(grid_for(static_cast<int64_t>(d) * d, 256)), dim3(256), 0, nullptr, Y.p, d,
                       Rm.p);
    check_launch("k_embed_real");
    DevBuf<int> dn(1);
    dn.upload(&m, 1);
    // jacobi_symmetric_all(R, max(5000, 150 m), 1e-13) (bench_ftlm_nmu_4x4_k.cpp:182-230, 324)
    launch_jacobi(1, m, nullptr, nullptr, nullptr, dn.p, nullptr, 1, std::max(5000, 150 * m), 1e-13, 0, scratch.p,
                  static_cast<int64_t>(m) * m, lam.p, nullptr, m, Rm.p);
    DevBuf<double> dbt(static_cast<size_t>(nb)), dl(static_cast<size_t>(nb));
    dbt.upload(betas, static_cast<size_t>(nb));
    hipLaunchKernelGGL(k_ed_lnz, dim3(grid_for(nb, 64)), dim3(64), 0, nullptr, lam.p, m, dbt.p, nb, dl.p);
    check_launch("k_ed_lnz");
    FTLM_HIP_CHECK(hipMemcpy(lnZ, dl.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
    if (evals) FTLM_HIP_CHECK(hipMemcpy(evals, lam.p, sizeof(double) * m, hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_dense_jacobi(const double* A, int n, int max_rotations, double tol, double* diag, double* lo,
                          double* hi) {
  return guarded([&] {
    FTLM_REQUIRE(A != nullptr && n >= 1 && max_rotations >= 1, FTLM_EINVAL, "bad arguments");
    FTLM_REQUIRE(jacobi_shmem(n, n <= kJacLdsMaxN) <= 60 * 1024 && n <= 2048, FTLM_EUNSUPPORTED,
                 "matrix too large for the Jacobi kernel");
    const int64_t nn = static_cast<int64_t>(n) * n;
    DevBuf<double> dA(static_cast<size_t>(nn)), scratch(static_cast<size_t>(nn)), lam(static_cast<size_t>(n));
    dA.upload(A, static_cast<size_t>(nn));
    DevBuf<int> dn(1);
    dn.upload(&n, 1);
    launch_jacobi(1, n, nullptr, nullptr, nullptr, dn.p, nullptr, 1, max_rotations, tol, 0, scratch.p, nn, lam.p,
                  nullptr, n, dA.p);
    std::vector<double> d(static_cast<size_t>(n));
    FTLM_HIP_CHECK(hipMemcpy(d.data(), lam.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (diag) std::copy(d.begin(), d.end(), diag);
    if (lo) *lo = *std::min_element(d.begin(), d.end());
    if (hi) *hi = *std::max_element(d.begin(), d.end());
  });
}

int ftlm_hip_tridiagonal_extrema(const double* alpha, int n, const double* beta, int nb, double* lo, double* hi) {
  return guarded([&] {
    FTLM_REQUIRE(lo != nullptr && hi != nullptr, FTLM_EINVAL, "null output");
    if (n <= 0) {
      *lo = 0.0;
      *hi = 0.0;
      return;
    }
    FTLM_REQUIRE(alpha != nullptr && (nb <= 0 || beta != nullptr), FTLM_EINVAL, "null alpha / beta");
    DevBuf<double> da(static_cast<size_t>(n)), db(static_cast<size_t>(std::max(n, 1))), dr(static_cast<size_t>(n));
    std::vector<double> hb(static_cast<size_t>(std::max(n, 1)), 0.0);
    for (int i = 0; i < std::min(nb, n); ++i) hb[i] = beta[i];
    da.upload(alpha, static_cast<size_t>(n));
    db.upload(hb);
    DevBuf<int> dn(1), dnb(1);
    const int nbv = std::max(0, nb);
    dn.upload(&n, 1);
    dnb.upload(&nbv, 1);
    quadrature_device(da.p, db.p, dn.p, dnb.p, 1, 1, n, FTLM_QUAD_REFERENCE, 120, 1e-15, 0, dr.p, nullptr, n, nullptr);
    std::vector<double> r(static_cast<size_t>(n));
    FTLM_HIP_CHECK(hipMemcpy(r.data(), dr.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    double a = r[0], b = r[0];
    for (double v : r) {
      a = std::min(a, v);
      b = std::max(b, v);
    }
    *lo = a;
    *hi = b;
  });
}

int ftlm_hip_quadrature_ritz(const double* alpha, int n, const double* beta, int nb, int quad, double* ritz,
                             double* weight) {
  return guarded([&] {
    FTLM_REQUIRE(n >= 0 && ritz != nullptr && weight != nullptr, FTLM_EINVAL, "bad arguments");
    if (n == 0) return;
    if (n == 1) {
      ritz[0] = alpha[0];
      weight[0] = 1.0;
      return;
    }
    DevBuf<double> da(static_cast<size_t>(n)), db(static_cast<size_t>(n)), dr(static_cast<size_t>(n)),
        dw(static_cast<size_t>(n));
    std::vector<double> hb(static_cast<size_t>(n), 0.0);
    for (int i = 0; i < std::min(nb, n); ++i) hb[i] = beta[i];
    da.upload(alpha, static_cast<size_t>(n));
    db.upload(hb);
    DevBuf<int> dn(1), dnb(1);
    const int nbv = std::max(0, nb);
    dn.upload(&n, 1);
    dnb.upload(&nbv, 1);
    quadrature_device(da.p, db.p, dn.p, dnb.p, 1, 1, n, quad, 0, 1e-14, 1, dr.p, dw.p, n, nullptr);
    FTLM_HIP_CHECK(hipMemcpy(ritz, dr.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    FTLM_HIP_CHECK(hipMemcpy(weight, dw.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  });
}

int ftlm_hip_ftlm_thermo_noalloc(const double* ritz, const double* weight, const int* used, int n_ops, int R, int M,
                                 const int64_t* dims, const int* N, const double* betas, int nb, const double* mus,
                                 int nmu, double* lnZ_op, double* E_op, double* lnZ_gc, double* N_gc, double* E_gc) {
  return guarded([&] {
    FTLM_REQUIRE(n_ops >= 1 && R >= 1 && M >= 1 && nb >= 1 && nmu >= 0, FTLM_EINVAL, "bad sizes");
    FTLM_REQUIRE(ritz && weight && used && dims && betas, FTLM_EINVAL, "null input");
    FTLM_REQUIRE(nmu == 0 || (N != nullptr && mus != nullptr), FTLM_EINVAL, "mu grid needs N and mus");
    int dev = 0;
    FTLM_HIP_CHECK(hipGetDevice(&dev));
    // every piece is fully written (inputs by the copies, outputs by the kernels) before it is
    // read, so the workspace left by an earlier call needs no clearing
    ScratchCarve sc{thermo_scratch(dev)};
    const size_t rk = static_cast<size_t>(n_ops) * R * M;
    const size_t ob = static_cast<size_t>(n_ops) * nb, gb = static_cast<size_t>(nb) * std::max(nmu, 1);
    double* dr = sc.take<double>(rk);
    double* dw = sc.take<double>(rk);
    int* du = sc.take<int>(static_cast<size_t>(n_ops) * R);
    int64_t* dd = sc.take<int64_t>(static_cast<size_t>(n_ops));
    double* dbt = sc.take<double>(static_cast<size_t>(nb));
    double* dmu = sc.take<double>(static_cast<size_t>(std::max(nmu, 1)));
    int* dN = sc.take<int>(static_cast<size_t>(n_ops));
    double* dl = sc.take<double>(ob);
    double* de = sc.take<double>(ob);
    double* dlg = sc.take<double>(gb);
    double* dng = sc.take<double>(gb);
    double* deg = sc.take<double>(gb);
    FTLM_HIP_CHECK(hipMemcpy(dr, ritz, rk * sizeof(double), hipMemcpyHostToDevice));
    FTLM_HIP_CHECK(hipMemcpy(dw, weight, rk * sizeof(double), hipMemcpyHostToDevice));
    FTLM_HIP_CHECK(hipMemcpy(du, used, static_cast<size_t>(n_ops) * R * sizeof(int), hipMemcpyHostToDevice));
    FTLM_HIP_CHECK(hipMemcpy(dd, dims, static_cast<size_t>(n_ops) * sizeof(int64_t), hipMemcpyHostToDevice));
    FTLM_HIP_CHECK(hipMemcpy(dbt, betas, static_cast<size_t>(nb) * sizeof(double), hipMemcpyHostToDevice));
    if (nmu > 0) {
      FTLM_HIP_CHECK(hipMemcpy(dmu, mus, static_cast<size_t>(nmu) * sizeof(double), hipMemcpyHostToDevice));
      FTLM_HIP_CHECK(hipMemcpy(dN, N, static_cast<size_t>(n_ops) * sizeof(int), hipMemcpyHostToDevice));
    }
    thermo_device(dr, dw, du, n_ops, R, M, dd, nmu > 0 ? dN : nullptr, dbt, nb, dmu, nmu, dl, de, dlg, dng, deg,
                  nullptr);
    FTLM_HIP_CHECK(hipDeviceSynchronize());
    if (lnZ_op) FTLM_HIP_CHECK(hipMemcpy(lnZ_op, dl, ob * sizeof(double), hipMemcpyDeviceToHost));
    if (E_op) FTLM_HIP_CHECK(hipMemcpy(E_op, de, ob * sizeof(double), hipMemcpyDeviceToHost));
    if (nmu > 0) {
      const size_t g = static_cast<size_t>(nb) * nmu * sizeof(double);
      if (lnZ_gc) FTLM_HIP_CHECK(hipMemcpy(lnZ_gc, dlg, g, hipMemcpyDeviceToHost));
      if (N_gc) FTLM_HIP_CHECK(hipMemcpy(N_gc, dng, g, hipMemcpyDeviceToHost));
      if (E_gc) FTLM_HIP_CHECK(hipMemcpy(E_gc, deg, g, hipMemcpyDeviceToHost));
    }
  });
}

int ftlm_hip_ftlm_thermo(const double* ritz, const double* weight, const int* used, int n_ops, int R, int M,
                         const int64_t* dims, const int* N, const double* betas, int nb, const double* mus, int nmu,
                         double* lnZ_op, double* E_op, double* lnZ_gc, double* N_gc, double* E_gc) {
  return guarded([&] {
    FTLM_REQUIRE(n_ops >= 1 && R >= 1 && M >= 1 && nb >= 1 && nmu >= 0, FTLM_EINVAL, "bad sizes");
    FTLM_REQUIRE(ritz && weight && used && dims && betas, FTLM_EINVAL, "null input");
    const size_t rk = static_cast<size_t>(n_ops) * R * M;
    DevBuf<double> dr(rk), dw(rk), dbt(static_cast<size_t>(nb)), dmu(static_cast<size_t>(std::max(nmu, 1)));
    DevBuf<int> du(static_cast<size_t>(n_ops) * R), dN(static_cast<size_t>(n_ops));
    DevBuf<int64_t> dd(static_cast<size_t>(n_ops));
    dr.upload(ritz, rk);
    dw.upload(weight, rk);
    du.upload(used, static_cast<size_t>(n_ops) * R);
    dd.upload(dims, static_cast<size_t>(n_ops));
    dbt.upload(betas, static_cast<size_t>(nb));
    if (nmu > 0) {
      FTLM_REQUIRE(N != nullptr && mus != nullptr, FTLM_EINVAL, "mu grid needs N and mus");
      dmu.upload(mus, static_cast<size_t>(nmu));
      dN.upload(N, static_cast<size_t>(n_ops));
    }
    const size_t ob = static_cast<size_t>(n_ops) * nb, gb = static_cast<size_t>(nb) * std::max(nmu, 1);
    DevBuf<double> dl(ob), de(ob), dlg(gb), dng(gb), deg(gb);
    thermo_device(dr.p, dw.p, du.p, n_ops, R, M, dd.p, nmu > 0 ? dN.p : nullptr, dbt.p, nb, dmu.p, nmu, dl.p, de.p,
                  dlg.p, dng.p, deg.p, nullptr);
    FTLM_HIP_CHECK(hipDeviceSynchronize());
    if (lnZ_op) FTLM_HIP_CHECK(hipMemcpy(lnZ_op, dl.p, ob * sizeof(double), hipMemcpyDeviceToHost));
    if (E_op) FTLM_HIP_CHECK(hipMemcpy(E_op, de.p, ob * sizeof(double), hipMemcpyDeviceToHost));
    if (nmu > 0) {
      const size_t g = static_cast<size_t>(nb) * nmu * sizeof(double);
      if (lnZ_gc) FTLM_HIP_CHECK(hipMemcpy(lnZ_gc, dlg.p, g, hipMemcpyDeviceToHost));
      if (N_gc) FTLM_HIP_CHECK(hipMemcpy(N_gc, dng.p, g, hipMemcpyDeviceToHost));
      if (E_gc) FTLM_HIP_CHECK(hipMemcpy(E_gc, deg.p, g, hipMemcpyDeviceToHost));
    }
  });
}

}  // extern "C"
