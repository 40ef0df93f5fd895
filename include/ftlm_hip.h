/* ftlm_hip.h -- C ABI of the MI355X (gfx950) FTLM hot-path library libftlm_hip.so.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No exceptions cross this
 * boundary; the C++ header-compatible layer (ftlm_extended_hubbard_amd/csrc/host) maps status
 * codes back onto the exceptions the reference throws.  Every entry point names the reference
 * interface it replaces (paths relative to the reference checkout).
 *
 * Vector layout.  A block of R complex vectors over a sector of dimension dim = D_up * D_dn is
 * stored element-major with the vector index innermost:  v[(iu * D_dn + id) * R + r], where
 * iu / id are the lexicographic ranks of the up / down strings.  For R = 1 this is exactly the
 * reference's x[iu * D_dn + id] (src/fock_basis.cpp:33-41).  Complex values are interleaved
 * (re, im) pairs of double (FTLM_C128) or float (FTLM_C64).
 *
 * Streams are hipStream_t passed as void* (NULL = the default stream).  Device pointers are
 * hipMalloc'd memory owned by the caller.  All entry points are thread-safe for distinct handles.
 */
#ifndef FTLM_HIP_H
#define FTLM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference exception types in parentheses) ---- */
#define FTLM_OK 0
#define FTLM_EINVAL 1       /* std::invalid_argument (src/fock_basis.cpp:19-24) */
#define FTLM_ERANGE 2       /* std::out_of_range (FockBasis::up_mask .at(), fock_basis.hpp:20-21) */
#define FTLM_EHIP 3         /* std::runtime_error: HIP runtime failure / no device */
#define FTLM_ENOMEM 4       /* std::bad_alloc */
#define FTLM_EUNSUPPORTED 5 /* std::invalid_argument: outside this library's limits */

/* ---- Krylov storage precision ---- */
#define FTLM_C128 0 /* complex<double> */
#define FTLM_C64 1  /* complex<float> storage; dots, norms, alpha/beta accumulate in FP64 */

/* ---- start vectors ---- */
#define FTLM_START_GAUSS_MT 0 /* reference: std::mt19937(seed) + normal_distribution, (re,im) per
                                 index, src/lanczos.cpp:140-175 (generated on the host) */
#define FTLM_START_PHASE 1    /* random phase e^{i theta}, theta ~ U[0,2pi), counter-based device RNG */

/* ---- Gauss-quadrature eigensolver ---- */
#define FTLM_QUAD_REFERENCE 0 /* capped max-pivot Jacobi, max(3000,120n) rotations at 1e-14
                                 (src/ftlm_thermo.cpp:17-146) */
#define FTLM_QUAD_CONVERGED 1 /* implicit-shift QL on the tridiagonal, run to convergence */

/* HubbardParams (include/ftlm/hubbard_params.hpp:6-15). */
typedef struct ftlm_params {
  int Lx, Ly;
  double t, tp, U, V, phi_x, phi_y;
} ftlm_params_t;

/* SpinfulHopping: H += coeff * c^dag_{to,spin} c_{from,spin} (include/ftlm/lattice.hpp:11-16). */
typedef struct ftlm_hop {
  int from, to, spin;
  double re, im;
} ftlm_hop_t;

/* NearestPair (include/ftlm/lattice.hpp:19-22). */
typedef struct ftlm_pair {
  int i, j;
} ftlm_pair_t;

/* A Hermitian operator resident on the device: a full (N_up, N_down) sector or an orbit k-block. */
typedef struct ftlm_op ftlm_op_t;

/* ------------------------------------------------------------------ library */
const char* ftlm_hip_version(void);    /* version_string (src/version.cpp:5) */
const char* ftlm_hip_last_error(void); /* message of this thread's last failing call */
int ftlm_hip_device_count(int* count);

/* nsites (src/hubbard_params.cpp:5). */
int ftlm_hip_nsites(const ftlm_params_t* p);

/* build_hubbard_geometry (src/lattice.cpp:38-110): same entries, same order, including the
 * duplicate bonds of small lattices.  n_hops / n_pairs receive the required sizes; returns
 * FTLM_ERANGE when a capacity is too small. */
int ftlm_hip_build_geometry(const ftlm_params_t* p, ftlm_hop_t* hops, int hop_cap, int* n_hops,
                            ftlm_pair_t* pairs, int pair_cap, int* n_pairs);

/* ------------------------------------------------------------------ full sectors */
/* FockBasis(n_sites, n_up, n_down) (src/fock_basis.cpp:17-42) + the operator data of
 * apply_extended_hubbard (src/hubbard_hamiltonian.cpp:51-108): enumerates the sector on the
 * device and builds the per-spin pull-hop tables.  Limits: n_sites <= 32, each spin's string
 * count < 2^20, at most 510 distinct directed bonds per spin. */
int ftlm_hip_sector_create(int n_sites, int n_up, int n_dn, const ftlm_hop_t* hops, int n_hops,
                           const ftlm_pair_t* pairs, int n_pairs, double U, double V, ftlm_op_t** out);
int ftlm_hip_op_destroy(ftlm_op_t* op);

/* FockBasis::dim / n_up / n_down; for k-blocks dim_up = dim_dn = 0. */
int ftlm_hip_op_dim(const ftlm_op_t* op, int64_t* dim, int64_t* dim_up, int64_t* dim_dn);

/* FockBasis::up_mask / down_mask (include/ftlm/fock_basis.hpp:20-21) for indices
 * [first, first+count), copied to host arrays.  FTLM_ERANGE outside [0, dim). */
int ftlm_hip_sector_masks(const ftlm_op_t* op, int64_t first, int64_t count, uint64_t* up, uint64_t* dn);

/* The per-spin string tables behind FockBasis::up_mask / down_mask (src/fock_basis.cpp:26-41):
 * spin 0 writes the D_up up strings, spin 1 the D_dn down strings, in lexicographic order (the
 * state index is i_up * D_dn + i_dn). */
int ftlm_hip_sector_strings(const ftlm_op_t* op, int spin, uint32_t* out);

/* FockBasis::index_of (src/fock_basis.cpp:44-51), batched, ranked on the device by the
 * combinatorial number system; -1 for states outside the sector.  Host arrays. */
int ftlm_hip_index_of(const ftlm_op_t* op, const uint64_t* up, const uint64_t* dn, int64_t count,
                      int64_t* out);

/* ------------------------------------------------------------------ H * v */
/* y = H x for a block of R vectors (layout above), device pointers, asynchronous on `stream`.
 * Replaces apply_extended_hubbard (src/hubbard_hamiltonian.cpp:51-108) and apply_orbit_k_block
 * (src/orbit_k_block_basis.cpp:151-220).  y is overwritten.  R in [1, 32].  Applies of one
 * operator on one stream run in stream order; applies from different streams may overlap (each
 * stream gets its own tile queues, for up to 64 streams per operator; FTLM_EUNSUPPORTED beyond). */
int ftlm_hip_apply(const ftlm_op_t* op, const void* x, void* y, int R, int dtype, void* stream);

/* ftlm_hip_apply with the row pass (down bonds) and the column pass (up bonds, diagonal) timed
 * separately by HIP events recorded on `stream`; synchronises the stream.  No counterpart in the
 * reference: instrumentation for roofline reports.  Full-sector operators only; a single-pass
 * (small) sector reports its whole time as *ms_row. */
int ftlm_hip_apply_pass_times(const ftlm_op_t* op, const void* x, void* y, int R, int dtype, void* stream,
                              double* ms_row, double* ms_col);

/* H*V launch-shape autotuner (no counterpart in the reference, whose apply is one scalar loop).
 * A shape holds, for one (R, dtype) of one operator:
 *   rb    vectors per launch (the block of R is applied in launches of at most rb vectors);
 *   bpx   workgroups per XCD in each pass of a two-pass sector apply;
 *   wlog  log2 of the column pass's stripe width (two-pass sectors);
 *   grid  workgroups of a single-pass apply (small sectors, momentum blocks).
 * 0 (wlog: -1) means the library default.  ftlm_hip_autotune times ftlm_hip_apply on scratch
 * blocks for candidate shapes (coordinate descent: rb, then bpx or grid, then wlog; `reps` timed
 * applies each, median, after one warm-up), keeps the fastest in the operator -- every later
 * apply, Lanczos step and FTLM run of the operator with that R and dtype uses it -- and reports
 * it with its time, the default shape's time and the number of shapes tried.  set / get store and
 * read a shape directly (to restore a cached tuning).  Results never depend on the shape beyond
 * the summation order of the dot-product partials. */
typedef struct ftlm_hv_shape {
  int rb, bpx, wlog, grid;
  double ms;         /* median time of one apply with this shape (autotune only) */
  double ms_default; /* the same for the default shape (autotune only) */
  int n_tried;       /* shapes timed (autotune only) */
} ftlm_hv_shape_t;
int ftlm_hip_autotune(ftlm_op_t* op, int R, int dtype, int reps, void* stream, ftlm_hv_shape_t* out);
int ftlm_hip_set_shape(ftlm_op_t* op, int R, int dtype, const ftlm_hv_shape_t* shape);
int ftlm_hip_get_shape(const ftlm_op_t* op, int R, int dtype, ftlm_hv_shape_t* shape);

/* Drop-in for apply_extended_hubbard(p, basis, hops, pairs, x, y) on host complex<double> arrays
 * of length dim: copies in, applies, copies out (synchronous). */
int ftlm_hip_apply_host(const ftlm_op_t* op, const double* x, double* y);

/* ------------------------------------------------------------------ Lanczos */
/* lanczos_tridiagonal (src/lanczos.cpp:125-209): 3-vector Hermitian Lanczos without
 * reorthogonalisation, device resident, FP64 dots/norms.  q0 (host complex<double>, length dim)
 * may be NULL: then the reference start vector mt19937(seed)+normal is drawn.  alpha and beta
 * must hold max_steps entries; *n_beta = len(beta); *used = steps used (0 on failure). */
int ftlm_hip_lanczos(const ftlm_op_t* op, int max_steps, unsigned seed, const double* q0, double* alpha,
                     double* beta, int* n_beta, int* used);

/* lanczos_extrema (src/lanczos.cpp:211-231): Ritz min/max from the 120-rotation capped Jacobi. */
int ftlm_hip_lanczos_extrema(const ftlm_op_t* op, int max_steps, unsigned seed, double* min_eval,
                             double* max_eval, int* steps_used);

/* The capped max-pivot Jacobi of the reference's dense checks -- jacobi_min_max
 * (benchmarks/bench_2x2.cpp:32-95, 1600 rotations at 1e-15) and sorted_eigenvalues_jacobi
 * (benchmarks/bench_2x2_nmu.cpp:75-126, 12000 rotations at 1e-14) -- on the device for a real
 * symmetric n x n host matrix (row-major): at most max_rotations rotations, stopping when the
 * largest off-diagonal magnitude is below tol.  diag (n, unsorted, may be NULL) receives the final
 * diagonal, lo / hi (may be NULL) its extremes. */
int ftlm_hip_dense_jacobi(const double* A, int n, int max_rotations, double tol, double* diag, double* lo,
                          double* hi);

/* tridiagonal_extrema (src/lanczos.cpp:106-123), evaluated on the device. */
int ftlm_hip_tridiagonal_extrema(const double* alpha, int n, const double* beta, int nb, double* lo,
                                 double* hi);

/* ------------------------------------------------------------------ FTLM */
typedef struct ftlm_ftlm_opts {
  int R;              /* random start vectors per operator (FtlmParams::n_random) */
  int M;              /* Lanczos steps (FtlmParams::lanczos_steps) */
  unsigned seed;      /* FtlmParams::seed; sample r uses seed + r*100003 (src/ftlm_thermo.cpp:189) */
  int start;          /* FTLM_START_* */
  int dtype;          /* FTLM_C128 / FTLM_C64 Krylov storage */
  int quad;           /* FTLM_QUAD_* */
  int clamp_M_to_dim; /* 1: M_op = min(M, dim) */
  int real_variant;   /* 1: ftlm_log_partition_real semantics (Re/Im applied separately,
                         src/ftlm_thermo.cpp:150-211) -- identical for real H */
  int path;           /* FTLM_PATH_*: how the recurrences of several operators are scheduled */
  int reorth;         /* FTLM_REORTH_*; the reference never reorthogonalises (src/lanczos.cpp:180-206) */
} ftlm_ftlm_opts_t;

/* Reorthogonalisation of the Lanczos vectors (no counterpart in the reference).  Every mode other
 * than NONE keeps all M+1 Krylov blocks resident in device memory ((M+1) * dim * R * 16 bytes at
 * complex128, 8 at complex64) and runs on the per-operator paths (AUTO never picks the persistent
 * launch for it).  Overlaps and norms are FP64 in every dtype.
 *   FULL    classical Gram-Schmidt against every earlier vector, each step, with a second pass
 *           wherever the first one cut the norm below 1/sqrt(2) of its value (DGKS criterion);
 *   PARTIAL Simon's omega recurrence estimates |<q_j|q_k+1>| from alpha/beta on the device and
 *           reorthogonalises (as FULL, against every earlier vector) only at steps where the
 *           estimate exceeds sqrt(eps) and at the step after; eps is the storage precision's unit
 *           roundoff, so the basis stays semi-orthogonal. */
#define FTLM_REORTH_NONE 0
#define FTLM_REORTH_FULL 1
#define FTLM_REORTH_PARTIAL 2

#define FTLM_PATH_AUTO 0       /* the full sectors of dim <= 16384 in one persistent launch (when
                                  there are at least two); momentum blocks of one orbit graph (at
                                  least two, random-phase starts) as in FTLM_PATH_STACK; every
                                  other operator per operator */
#define FTLM_PATH_PER_OP 1     /* per operator, R recurrences per block H*V */
#define FTLM_PATH_PERSISTENT 2 /* every (sector, sample) recurrence in one persistent launch (sectors
                                  of dim <= 262144) */
#define FTLM_PATH_GRAPH 3      /* per operator, the whole M-step recurrence captured as one hipGraph
                                  and launched once (host-callback operators run eagerly) */
#define FTLM_PATH_STACK 4      /* the momentum blocks that share one orbit graph as momentum stacks
                                  (ftlm_hip_kstack_create): one recurrence group and one apply per
                                  step for all of them; every other operator per operator */

/* Runs R independent Lanczos recurrences per operator, all on the device, and diagonalises every
 * tridiagonal with the chosen quadrature eigensolver.  Outputs (host):
 *   ritz[n_ops][R][M]   Ritz values (first used entries valid)
 *   weight[n_ops][R][M] squared first eigenvector components
 *   used[n_ops][R]      Lanczos steps used (0: sample skipped)
 * alpha/beta[n_ops][R][M] may be NULL.  Operators are processed in order; each one's R samples
 * run as one block recurrence (block H*V). */
int ftlm_hip_ftlm_ritz(ftlm_op_t* const* ops, int n_ops, const ftlm_ftlm_opts_t* opts, double* ritz,
                       double* weight, int* used, double* alpha, double* beta, void* stream);

/* ftlm_hip_ftlm_ritz with a seed per operator: operator io's sample r uses op_seeds[io] + r*100003
 * (opts->seed is not used).  This is the reference's per-K seed schedule of its momentum loop
 * (seed + sector*257 + ky*17 + kx, benchmarks/bench_ftlm_nmu_4x4_k.cpp:1077-1078) in one call, so
 * the blocks of a sector can share momentum stacks; with equal seeds it is ftlm_hip_ftlm_ritz. */
int ftlm_hip_ftlm_ritz_seeded(ftlm_op_t* const* ops, int n_ops, const ftlm_ftlm_opts_t* opts,
                              const unsigned* op_seeds, double* ritz, double* weight, int* used, double* alpha,
                              double* beta, void* stream);

/* Per-operator FTLM thermodynamics from Ritz data, on the device:
 *   lnZ_op[n_ops][nb] = ln dim + log-mean-exp_r ln sum_k w e^{-beta lambda}
 *                      (ftlm_log_partition_complex, src/ftlm_thermo.cpp:226-255)
 *   E_op[n_ops][nb]   = sum_rk w lambda e^{-beta lambda} / sum_rk w e^{-beta lambda}
 * and the grand-canonical mixture over operators with particle numbers N[n_ops]
 * (benchmarks/bench_ftlm_nmu_rect.cpp:186-219), for every (beta, mu):
 *   lnZ_gc[nb][nmu], N_gc[nb][nmu] (<N>), E_gc[nb][nmu] (<H>, mu excluded).
 * Any output pointer may be NULL. */
int ftlm_hip_ftlm_thermo(const double* ritz, const double* weight, const int* used, int n_ops, int R, int M,
                         const int64_t* dims, const int* N, const double* betas, int nb, const double* mus,
                         int nmu, double* lnZ_op, double* E_op, double* lnZ_gc, double* N_gc, double* E_gc);

/* ftlm_hip_ftlm_thermo without any device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (harnesses that forbid hipMalloc, code paths that
 * must not allocate).  Inputs, outputs and the per-operator intermediate sums are staged in a
 * file-scope device workspace of 16 MB per device (held until the process ends); requests
 * that need more return FTLM_EUNSUPPORTED.  Same arguments and results as ftlm_hip_ftlm_thermo, on
 * the current device, synchronous.  Every call on a device shares that device's workspace:
 * calls that overlap on one device (other host threads) race. */
int ftlm_hip_ftlm_thermo_noalloc(const double* ritz, const double* weight, const int* used, int n_ops, int R, int M,
                                 const int64_t* dims, const int* N, const double* betas, int nb, const double* mus,
                                 int nmu, double* lnZ_op, double* E_op, double* lnZ_gc, double* N_gc, double* E_gc);

/* ftlm_log_partition_complex / _real (src/ftlm_thermo.cpp:150-256) for a list of betas (the
 * Lanczos data are beta independent).  opts->start = FTLM_START_GAUSS_MT reproduces the
 * reference start vectors.  dim == 1 returns -beta*H00; dim <= 0 returns -inf. */
int ftlm_hip_ftlm_log_partition(const ftlm_op_t* op, const double* betas, int nb, const ftlm_ftlm_opts_t* opts,
                                double* lnZ);

/* Capped-Jacobi Gauss quadrature of one tridiagonal (log_lanczos_tridiagonal_quadrature_exp,
 * src/ftlm_thermo.cpp:83-146), on the device: Ritz values and weights of n entries. */
int ftlm_hip_quadrature_ritz(const double* alpha, int n, const double* beta, int nb, int quad, double* ritz,
                             double* weight);

/* ------------------------------------------------------------------ step-level Krylov sessions */
/* The three-vector workspace of the reference (LanczosComplexWorkspace, include/ftlm/lanczos.hpp:13-35)
 * as a device object holding R recurrences: create draws the start block (FTLM_START_*; sample r
 * uses seed + r*100003), run advances n_steps (up to max_steps in total) asynchronously on
 * `stream`, result copies alpha/beta [max_steps][R], used [R] and alive [R] to the host
 * (synchronising); bytes reports the session's device memory footprint. */
typedef struct ftlm_krylov ftlm_krylov_t;
int ftlm_hip_krylov_create(const ftlm_op_t* op, int R, int max_steps, int dtype, int start, unsigned seed,
                           ftlm_krylov_t** out);
int ftlm_hip_krylov_run(ftlm_krylov_t* k, int n_steps, void* stream);
/* The same n_steps captured as one hipGraph (kernel nodes only) and launched once on a private
 * stream ordered after `stream`; results are identical to ftlm_hip_krylov_run.  No counterpart in
 * the reference. */
int ftlm_hip_krylov_run_graph(ftlm_krylov_t* k, int n_steps, void* stream);
int ftlm_hip_krylov_result(ftlm_krylov_t* k, double* alpha, double* beta, int* used, int* alive);
int ftlm_hip_krylov_bytes(const ftlm_krylov_t* k, int64_t* bytes);
int ftlm_hip_krylov_destroy(ftlm_krylov_t* k);

/* exact_log_partition_small_complex (benchmarks/bench_ftlm_nmu_4x4_k.cpp:249-336) on the device:
 * dense H of the operator (dim <= 1024), its real 2d x 2d embedding, the reference's capped
 * max-pivot Jacobi (max(5000, 150 m) rotations, 1e-13), ln Z = ln(1/2) + log-sum-exp(-beta lambda)
 * for each of nb betas.  dim <= 0 gives -inf, dim == 1 gives -beta H00.  evals (optional, 2*dim)
 * receives the Jacobi diagonal: every eigenvalue of H twice, unsorted. */
int ftlm_hip_exact_log_partition(const ftlm_op_t* op, const double* betas, int nb, double* lnZ, double* evals);

/* ------------------------------------------------------------------ derivatives */
/* d(coefficient)/d(theta) of every entry of build_hubbard_geometry (same order):
 * which = 0 (t), 1 (t'), 2 (phi_x), 3 (phi_y).  With U' and V' set to 0 this gives the hop part
 * of dH/dtheta; dH/dU and dH/dV are the sector operators with no hops and (U, V) = (1, 0) / (0, 1).
 * No counterpart in the reference (its parameters enter src/lattice.cpp:38-110). */
int ftlm_hip_build_geometry_derivative(const ftlm_params_t* p, int which, ftlm_hop_t* hops, int hop_cap,
                                       int* n_hops);

/* lnZ(beta) of the FTLM estimator (ftlm_log_partition_complex, src/ftlm_thermo.cpp:213-256, with
 * the converged Gauss rule) and its exact derivatives dlnZ/dtheta_p for npar parameters whose
 * dH/dtheta_p are given as operators of the same size (forward-mode differentiation of the
 * Lanczos recurrence and of the quadrature).  opts: R <= 32 samples (seed + r*100003, Gaussian or
 * random-phase starts), M <= 512 steps, complex128.  grad is [npar][nb].  No counterpart in the
 * reference, which has no derivatives. */
int ftlm_hip_ftlm_gradient(const ftlm_op_t* op, ftlm_op_t* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                           const double* betas, int nb, double* lnZ, double* grad);

/* Vector-Jacobian product of the same estimator (reverse mode): lnZ(beta_b) for nb betas and
 * grad_p = sum_b gw[b] dlnZ(beta_b)/dtheta_p, from one forward Lanczos pass that keeps
 * (q_{k-1}, q_k) every ckpt steps (0: round(sqrt(2M))), the quadrature's derivatives, and one
 * adjoint pass that recomputes each segment's Krylov vectors from its checkpoint.  The parameter
 * adjoints sum_k Re <w-bar_k| dH_p |q_k> come from one term-split pass over the operator's own
 * terms -- a full sector's hop lists or a momentum block's orbit terms, all derivative tables and
 * the U / V terms at once -- when every dH_p is a hop table over H's bonds plus U / V terms (mode
 * FTLM_TS_AUTO or FTLM_TS_FUSED), otherwise from one apply of each dH_p (FTLM_TS_GENERIC, and
 * callback operators).  The backward sweep applies H as H^dag:
 * momentum blocks whose hop coefficients are not translation invariant (flux on wrapped bonds)
 * are refused with FTLM_EUNSUPPORTED; callback operators must be Hermitian.  grad is [npar].  No
 * counterpart in the reference. */
#define FTLM_TS_AUTO 0
#define FTLM_TS_GENERIC 1
#define FTLM_TS_FUSED 2
int ftlm_hip_ftlm_vjp(const ftlm_op_t* op, ftlm_op_t* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                      const double* betas, int nb, const double* gw, int ckpt, int mode, double* lnZ, double* grad);

/* FTLM estimator of observables A_p (Hermitian operators of the same size, e.g. the derivative
 * operators: double occupancy, NN pair count, the hop part):
 *   <A>(beta) = sum_r sum_j e^{-beta lambda_j} <r|psi_j><psi_j|A|r> / sum_r sum_j e^{-beta lambda_j} |<r|psi_j>|^2
 * over the same R samples, M steps and converged Gauss rule as ftlm_hip_ftlm_gradient, with
 * <psi_j|A|r> = sum_k U_kj <q_k|A r> (one apply of A_p to the start block and one dot per step).
 * Also returns lnZ.  dim > 1, 1 <= nobs <= 16; obs is [nobs][nb].  No counterpart in the
 * reference, which estimates only lnZ (for kinetic energy, double occupancy, NN correlations, N). */
int ftlm_hip_ftlm_observables(const ftlm_op_t* op, ftlm_op_t* const* aops, int nobs, const ftlm_ftlm_opts_t* opts,
                              const double* betas, int nb, double* lnZ, double* obs);

/* The two halves of ftlm_hip_ftlm_vjp, for callers that choose the weights after the forward pass
 * (reverse-mode autodiff) or need several weighted rows (a Jacobian).  tape_forward runs the
 * forward pass, returns lnZ and keeps the checkpoints, scalars and quadrature derivatives on the
 * tape; tape_backward(gw) runs one adjoint sweep and may be called any number of times.  op and
 * dops must stay alive until tape_destroy.  No counterpart in the reference. */
typedef struct ftlm_tape ftlm_tape_t;
int ftlm_hip_ftlm_tape_forward(const ftlm_op_t* op, ftlm_op_t* const* dops, int npar, const ftlm_ftlm_opts_t* opts,
                               const double* betas, int nb, int ckpt, int mode, double* lnZ, ftlm_tape_t** out);
int ftlm_hip_ftlm_tape_backward(const ftlm_tape_t* tape, const double* gw, double* grad);
int ftlm_hip_ftlm_tape_destroy(ftlm_tape_t* tape);

/* ------------------------------------------------------------------ host-callback operators */
/* The operator argument of lanczos_tridiagonal / lanczos_extrema / ftlm_log_partition_complex
 * (std::function over host complex<double> arrays, include/ftlm/lanczos.hpp:47-63,
 * include/ftlm/ftlm_thermo.hpp:41-50) as an operator handle: `apply(x, y, ctx)` computes y = H x
 * on interleaved (re, im) host arrays of length dim.  The Lanczos / FTLM recurrences then run on
 * the device; each H*v copies one vector to the host and back.  complex128 only. */
typedef void (*ftlm_host_apply_fn)(const double* x, double* y, void* ctx);
int ftlm_hip_callback_op_create(int64_t dim, ftlm_host_apply_fn apply, void* ctx, ftlm_op_t** out);

/* ------------------------------------------------------------------ translations and k-blocks */
/* States are packed RawStates, (up << 16) | dn (include/ftlm/symmetry/packed_raw_state.hpp:12-14);
 * lattices have at most 16 sites; translation t = (dx, dy). */

/* translate_occupation_bits_rect (include/ftlm/symmetry/occup_bits_translate.hpp:10-25) on both
 * strings and fermionic_translation_sign_rect (include/ftlm/symmetry/fermionic_translation.hpp:32-54),
 * batched on the device; out and/or sign may be NULL. */
int ftlm_hip_translate_states(int Lx, int Ly, const uint32_t* states, int64_t count, int dx, int dy, uint32_t* out,
                              double* sign);

/* canonical_raw_state_rect (minimal pack over the orbit) and the size of
 * translation_stabilizer_rect (include/ftlm/symmetry/orbit.hpp:93-155), batched on the device. */
int ftlm_hip_canonical_states(int Lx, int Ly, const uint32_t* states, int64_t count, uint32_t* canon,
                              int* stab_size);

/* translation_bloch_phase(K, dx, dy) = exp(-2 pi i (kx dx / Lx + ky dy / Ly))
 * (include/ftlm/symmetry/momentum_sector.hpp:57-63). */
int ftlm_hip_bloch_phase(int Lx, int Ly, int kx, int ky, int dx, int dy, double* re, double* im);

/* OrbitKBlockBasis::build(FockBasis(Lx*Ly, n_up, n_dn), MomentumSector{kx, ky, Lx, Ly}, tol)
 * (src/orbit_k_block_basis.cpp:70-117) plus the operator data of apply_orbit_k_block
 * (src/orbit_k_block_basis.cpp:151-220): hops and pairs as build_hubbard_geometry returns them.
 * The orbit graph (representatives and the K-independent pull-form hop table) is built on the
 * device; pass another k-block of the same sector and Hamiltonian as `share` (or NULL) to reuse
 * its graph.  The result is an operator for ftlm_hip_apply / _lanczos / _ftlm_* of dimension
 * OrbitKBlockBasis::dim(). */
int ftlm_hip_kblock_create(int Lx, int Ly, int n_up, int n_dn, int kx, int ky, double tol, const ftlm_hop_t* hops,
                           int n_hops, const ftlm_pair_t* pairs, int n_pairs, double U, double V,
                           const ftlm_op_t* share, ftlm_op_t** out);

/* Momentum, number of translation orbits of the sector and stored hop terms of a k-block. */
int ftlm_hip_kblock_info(const ftlm_op_t* op, int* kx, int* ky, int64_t* n_orbits, int64_t* nnz);

/* OrbitKBlockBasis::entries() rows [first, first+count): rep (packed), orbit_size,
 * stabilizer_size, A_k as (re, im) pairs (include/ftlm/symmetry/orbit_k_block_basis.hpp:35-42);
 * any output may be NULL. */
int ftlm_hip_kblock_entries(const ftlm_op_t* op, int64_t first, int64_t count, uint32_t* rep, int* orbit_size,
                            int* stab_size, double* A);

/* OrbitKRepEntry::orbit_index of rows [first, first+count): the row's position in the sector's
 * pack-sorted orbit list (the reference indexes its own MomentumSectorMap, whose order differs). */
int ftlm_hip_kblock_orbit_index(const ftlm_op_t* op, int64_t first, int64_t count, int64_t* out);

/* OrbitKBlockBasis::reduced_index_of_rep (src/orbit_k_block_basis.cpp:119-125), batched; -1 when
 * the representative is not a row of the block. */
int ftlm_hip_kblock_index_of_rep(const ftlm_op_t* op, const uint32_t* reps, int64_t count, int64_t* out);

/* OrbitKBlockBasis::map_raw_to_rep (src/orbit_k_block_basis.cpp:127-149), batched: row, (dx, dy)
 * of the first translation (dx outer, dy inner) with T rep = s, and eta; index -1 (dx = dy = 0,
 * eta = 0) where the reference returns false. */
int ftlm_hip_kblock_map(const ftlm_op_t* op, const uint32_t* states, int64_t count, int64_t* index, int* dx, int* dy,
                        double* eta);

/* Momentum stack: the momentum blocks of one sector and Hamiltonian (created with share= so that
 * they use one orbit graph) applied side by side, replacing the reference's loop that builds and
 * runs apply_orbit_k_block once per K (benchmarks/bench_ftlm_nmu_4x4_k.cpp:801-1103 over
 * src/orbit_k_block_basis.cpp:151-220).  The stack's dim is the number of orbits of the graph; an
 * apply takes R = n_blocks * rk vectors laid out [orbit][block j][rk]: slot (j, r) of orbit g holds
 * block j's entry at the row whose ftlm_hip_kblock_orbit_index is g, and the slots of blocks that g is not a row
 * of hold 0 and are written as 0.  Each block's slots of y equal that block's own apply.  The
 * blocks must outlive the stack (it borrows their row maps); at most 16 blocks.  ftlm_hip_apply,
 * ftlm_hip_krylov_* (random-phase starts only: slot (j, r) starts as block j's sample r) and
 * ftlm_hip_ftlm_ritz take it as an operator. */
int ftlm_hip_kstack_create(ftlm_op_t* const* blocks, int n_blocks, ftlm_op_t** out);
/* Block j's vectors xb[dim_j][rk] into slots (j, 0..rk) of a stacked block xs[n_orbits][n_blocks*rk]
 * (scatter writes 0 into the orbits that are not rows of block j), and back (gather). */
int ftlm_hip_kstack_scatter(const ftlm_op_t* stack, int j, const void* xb, int rk, int dtype, void* xs,
                            void* stream);
int ftlm_hip_kstack_gather(const ftlm_op_t* stack, int j, const void* xs, int rk, int dtype, void* xb,
                           void* stream);

/* HubbardMomentumBlock (src/hubbard_momentum_action.cpp:37-89) with the Gram basis of
 * MomentumPhiGramBasis (src/orbit_bloch_phi.cpp:226-355): the momentum-K block of `sector` as
 * y = Phi^dag H Phi x, applied by lifting into the full sector, the sector apply and a projection
 * (each apply costs one full-sector apply, as in the reference).  `kblock` is a momentum block of
 * the same lattice, particle numbers and K (ftlm_hip_kblock_create); it may be destroyed after
 * this call, `sector` must outlive the result.  The block keeps the reference's subspace: the
 * Bloch sums of the orbits whose stabiliser has a trivial bosonic character at K
 * (momentum_phi_seeds, include/ftlm/symmetry/k_basis.hpp:66-84) that survive the Gram rank cut,
 * i.e. those whose fermionic translation signs are trivial on the stabiliser too; orbits with an
 * odd stabiliser sign drop out of every K, as in the reference.  The columns are one normalised
 * Bloch sum per orbit in pack order; the reference's own orthonormal basis of the same space comes
 * from LAPACK zheev, so block spectra and exact lnZ agree, FTLM samples agree in distribution.
 * The operator is Hermitian; graph capture and launch-shape tuning do not apply to it. */
int ftlm_hip_kgram_create(const ftlm_op_t* sector, const ftlm_op_t* kblock, ftlm_op_t** out);

#ifdef __cplusplus
}
#endif

#endif /* FTLM_HIP_H */
