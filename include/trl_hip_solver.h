/*
 * trl_hip_solver.h — C-ABI of libtrl_hip_solver.so: whole thick-restart
 * block-Lanczos solves on the MI355X backend, for callers that are not C++
 * (ctypes, the benchmark, the autograd wrapper).
 *
 * Each entry point constructs the trl::hip eigenproblem over caller-owned
 * device arrays, runs trl_amd::BlockLanczos (the restatement of the
 * reference driver include/trl/eigensolvers/lanczos.hh:22-312; parameters as
 * in include/trl/eigensolvers/params.hh:5-16) and returns the Ritz pairs.
 * The reference has no C-ABI; these entry points stand where its C++
 * callers construct BlockLanczos and call solve() (examples/openmp/csrtest.cc:
 * 100-103, tests/test_lanczos_convergence.hh:113-122).
 *
 * Status codes are those of trl_hip.h (0 ok, <0 TRL_HIP_E*, >0 hipError_t).
 */
#ifndef TRL_HIP_SOLVER_H
#define TRL_HIP_SOLVER_H

#include "trl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int b;                 /* block size: 1, 2, 4, 6, 8, 16                        */
  unsigned nev;          /* wanted eigenpairs             (params.hh:6)          */
  unsigned ncv;          /* basis size, multiple of b     (params.hh:7)          */
  unsigned max_restarts; /* restart cycles                (params.hh:8)          */
  double tol;            /* relative residual tolerance   (params.hh:9)          */
  int order;             /* 0 |lambda| descending (evp_base.hh:103), 1 ascending,
                            2 descending                                         */
  int reorth;            /* 0 ModifiedGS (reorthogonalization.hh:12-22),
                            1 block CGS2 (two multi-block Gram/update passes),
                            2 block CGS with the DGKS second-pass test          */
  int cholqr_passes;     /* 1 CholQR (evp_base.hh:39-71), 2 CholQR2, 0 default  */
  int lanes_per_row;     /* operator kernel variant (trl_hip_spmat::variant),
                            0 = default                                          */
  int spmm_tune;         /* 1: time the operator kernels (CSR variants, ELL) for
                            b at the start of the solve and use the fastest;
                            lanes_per_row is then ignored                        */
  int rr_dense;          /* 1: the Rayleigh-Ritz step always takes the dense
                            Householder eigensolver (for comparison); 0: the
                            structured one whenever T has the thick-restart
                            form (diagonal Ritz part, arrow, band)              */
} trl_hip_solver_params;

typedef struct {
  int converged;           /* EigensolverResult::converged  (params.hh:13)   */
  unsigned iterations;     /* EigensolverResult::iterations (params.hh:14)   */
  unsigned n_op_apply;     /* EigensolverResult::n_op_apply (params.hh:15)   */
  double seconds;          /* wall time of solve(), device work included     */
  double host_seconds;     /* of which the host Rayleigh-Ritz steps          */
  unsigned rr_steps;       /* Rayleigh-Ritz steps                            */
  unsigned rr_structured;  /* of which by the structured eigensolver         */
  unsigned graph_cycles;   /* restart cycles replayed from a captured graph   */
} trl_hip_solver_result;

/* Standard (mass = mass_inv = NULL) or generalized problem with a diagonal
 * mass M (both device arrays of the panel dtype, mass_inv = 1/M).
 * val: CSR values in the panel dtype.  V0: device n x b initial block.
 * evals: host, nev Ritz values in the chosen order.  X (optional): device,
 * ceil(nev/b) panels (n x b each) receiving the Ritz vectors.
 * resid (optional): host, nev true relative residuals
 * ||A x - lambda M x|| / ||A x|| evaluated on the device. */
int trl_hip_solve_csr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, const void* val, const void* mass,
                      const void* mass_inv, const trl_hip_solver_params* p, const void* V0, double* evals, void* X, double* resid,
                      trl_hip_solver_result* res);

typedef struct {
  double tol_fp32;           /* stopping tolerance of the FP32 phase                  */
  unsigned max_restarts_fp32;
  unsigned long long seed;   /* refine = 0: mixing of the FP32 Ritz vectors           */
  int refine;                /* 0: FP64 Lanczos from a mix of the FP32 Ritz vectors,
                                1: Chebyshev + Rayleigh-Ritz sweeps, 2: FP64
                                thick-restart cycles from the FP32 restart state */
  unsigned cheb_degree;      /* refine = 1: filter degree per refinement sweep        */
  unsigned max_refine;       /* refine = 1: sweeps before falling back to refine = 0  */
  int phase1_fp64;           /* 1: run phase 1 in FP64 too (the same schedule without
                                the precision change, for comparisons)                */
  unsigned refine_blocks;    /* refine = 1: Ritz blocks refined together (0: nev/b + 7);
                                the blocks beyond nev/b + 1 are a buffer that moves
                                the filter cutoff away from the wanted values         */
  unsigned fp32_sweeps;      /* refine = 1: at most this many filter sweeps in FP32
                                panels first (0: none) ...                            */
  double tol_fp32_refine;    /* ... stopping once the FP32 residuals are below this   */
  int degree_opt;            /* refine = 1: 1 gives each Ritz block its own filter
                                degree from the measured convergence rate of its
                                vectors (0 when converged) in every sweep, 2 only
                                in the FP64 sweeps, 0: cheb_degree on all         */
  int residual_filter;       /* refine = 1 with FP32 sweeps: 1 runs the FP64 sweeps'
                                filter in FP32, on E = p(A) x - p(theta) x driven
                                by the FP64 residual A x - theta x of each Ritz
                                pair (x <- p(theta) x + E in FP64); 0: FP64 filter */
  unsigned ncv_fp32;         /* phase 1's Krylov space (a multiple of b; 0: ncv); the
                                FP64 Lanczos fallback keeps ncv                   */
} trl_hip_mixed_params;

/* Mixed precision.  Phase 1: an FP32 thick-restart solve (FP32 basis, SpMM
 * and Gram products with compensated sums, FP64 coefficients, Cholesky and
 * Rayleigh-Ritz) to tol_fp32, keeping its leading Ritz blocks X.
 * Phase 2 (FP64), by mp->refine:
 *   1: (optionally first the same sweeps in FP32 panels, mp->fp32_sweeps)
 *      block refinement of all of X together: FP64 block CGS + CholQR2,
 *      Rayleigh-Ritz, and while a true residual exceeds p->tol, a scaled
 *      Chebyshev filter of degree cheb_degree (Zhou & Saad, J. Comput. Phys.
 *      2007) that damps the spectrum beyond the kept Ritz values, then
 *      Rayleigh-Ritz again; after max_refine sweeps it falls back to 0.
 *      With mp->residual_filter the filter of the FP64 sweeps runs in FP32
 *      panels: for a Ritz pair (theta, x) with residual r = A x - theta x,
 *      E_i = p_i(A) x - p_i(theta) x obeys the filter's three-term
 *      recurrence plus alpha_i p_{i-1}(theta) r, and |E_i| is of the order
 *      of the error of x, so FP32 rounding in E stays below the FP64
 *      tolerance; x <- p(theta) x + E is formed in FP64;
 *   0: an FP64 thick-restart solve started from V0 = X G, G a seeded
 *      Gaussian (nev/b+1)*b x b mixing matrix;
 *   2: FP64 thick-restart cycles continuing the FP32 solve's restart state
 *      (lanczos.hh:97-143): its k_r = nev/b + 1 Ritz blocks and the block
 *      after them, converted to FP64 and orthonormalised, their FP64
 *      Rayleigh quotient diagonalised into theta and the arrow, then the
 *      quasi-Lanczos step and the driver's cycles.  The FP64 cycles stop on
 *      the Ritz estimate; the part of A V outside the seeded blocks (the
 *      FP32 basis' rounding) stays in the true residuals, which resid
 *      reports.
 * Order must be ascending or descending for refine = 1.
 * val32/mass32/minv32 are FP32 copies of the FP64 arrays.  evals, X, resid
 * refer to the final FP64 result; res->seconds is the wall time of both
 * phases, res->iterations the FP64 sweeps or cycles; res32 (optional)
 * describes the FP32 phase. */
int trl_hip_solve_csr_mixed(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, const double* val64, const float* val32,
                            const double* mass64, const double* minv64, const float* mass32, const float* minv32, const trl_hip_solver_params* p,
                            const trl_hip_mixed_params* mp, const double* V0, double* evals, double* X, double* resid, trl_hip_solver_result* res,
                            trl_hip_solver_result* res32);

/* ---- generalized problems by shift-invert ----------------------------------- */

typedef struct {
  unsigned solves;              /* inner solves (one per operator apply)            */
  unsigned long long iterations; /* conjugate gradient iterations over all solves   */
  unsigned max_iterations;      /* the most any one solve took                      */
} trl_hip_inner_stats;

/* A x = lambda B x with sparse symmetric A and sparse SPD B by shift-invert:
 * the reference's CSRGeneralizedEVP driven by BlockLanczos
 * (examples/openmp/csrevp.hh:51-104, examples/openmp/gevp_test.cc:62-116).
 * The driver runs on (A - shift B)^{-1} B in the B-inner product; its Ritz
 * values mu = 1/(lambda - shift) (order as p->order, |mu| descending by
 * default: the lambda nearest the shift) are returned mapped back,
 * evals[i] = shift + 1/mu[i] (gevp_test.cc:135-141), mu optional.
 * A, B: host CSR arrays (full symmetric storage, columns ascending per row),
 * FP64.  The inner solve is the device PCG of trl_hip_pcg at relative
 * tolerance inner_tol (0: 1e-12), so A - shift B must be positive definite
 * (TRL_HIP_EINDEF otherwise; TRL_HIP_ENOCONV if an inner solve stalls).
 * p->reorth 0 = ModifiedGS through the B-dot, 1 or 2 = block CGS2 in the
 * B-inner product.  V0: device n x b.  X (optional, device): ceil(nev/b)
 * panels of B-orthonormal Ritz vectors.  resid (optional, host): nev true
 * residuals |A x - lambda B x| / |A x|.  inner (optional): PCG counts. */
int trl_hip_solve_csr_shift_invert(trl_hip_ctx* ctx, size_t n, const int32_t* a_rowptr, const int32_t* a_col, const double* a_val,
                                   const int32_t* b_rowptr, const int32_t* b_col, const double* b_val, double shift, double inner_tol,
                                   const trl_hip_solver_params* p, const double* V0, double* evals, double* mu, double* X, double* resid,
                                   trl_hip_solver_result* res, trl_hip_inner_stats* inner);

/* ---- many independent problems in lock-step (no reference counterpart) ---- */

typedef struct {
  int b;                 /* block size: 1, 2, 4, 8                              */
  unsigned nev, ncv, max_restarts;
  double tol;
  int order;             /* as trl_hip_solver_params::order                     */
  int cholqr_passes;     /* 0/1: Cholesky-QR as the reference, 2: CholQR2       */
  int tune;              /* 1: autotune the step kernel's tiles and format      */
  int use_graph;         /* 1: replay each cycle's steps from a hipGraph        */
  int reorth;            /* 0 or 1: two block CGS passes per step; 2: the second
                            pass only where the DGKS test asks for it          */
} trl_hip_batch_params;

typedef struct {
  double seconds;        /* wall time of the restart loop (device + host)      */
  double host_seconds;   /* of which the host Rayleigh-Ritz steps              */
  double setup_seconds;  /* layout upload, start blocks, autotuning            */
  unsigned cycles;       /* restart cycles of the batch (the slowest problem)  */
  unsigned graph_launches;
  int ntiles, format, rows_per_thread;
  double tuned_us;       /* the chosen variant's time for one full-width step  */
  double extend_seconds; /* of the loop: the replayed step graphs (device)     */
  double rotate_seconds; /*              the restart rotations                 */
  unsigned retiles;      /* tilings rebuilt as problems finished               */
} trl_hip_batch_stats;

/* nprob standard FP64 problems A_p x = lambda x (host CSR arrays, host start
 * blocks V0[p] of n[p] x b) solved together by the thick-restart block
 * Lanczos recurrence of trl_hip_solve_csr (block CGS2 reorthogonalisation),
 * advanced in lock-step: one persistent kernel launch per step of every
 * active problem, each cycle's steps replayed from a hipGraph, the
 * Rayleigh-Ritz steps on the host in parallel threads; converged problems
 * leave the batch.  evals: nprob x nev; resid (optional): true relative
 * residuals, nprob x nev; X (optional): host arrays of ceil(nev/b) x n[p] x b;
 * res[p].seconds is the time from the start of the loop to the end of the
 * cycle in which problem p converged. */
int trl_hip_batch_solve_csr(trl_hip_ctx* ctx, int nprob, const size_t* n, const int32_t* const* rowptr, const int32_t* const* col,
                            const double* const* val, const double* const* V0, const trl_hip_batch_params* p, double* evals, double* resid,
                            double* const* X, trl_hip_solver_result* res, trl_hip_batch_stats* stats);

/* Host FP64 symmetric eigensolver used by the Rayleigh-Ritz step:
 * A (n x n row-major) -> ascending w, eigenvectors in the columns of Z. */
int trl_hip_dense_syev(int n, const double* A, double* w, double* Z);
/* The same eigenvalues (all n, ascending, in w) and only the eigenvectors of
 * w[sel[0]] < w[sel[1]] < ... (k strictly increasing indices) in the columns
 * of the row-major n x k matrix Z: inverse iteration on the tridiagonal form
 * with Gram-Schmidt inside clusters, then the Householder back-transform.
 * The Rayleigh-Ritz step uses it for the k_r*b Ritz vectors the driver
 * needs. */
int trl_hip_dense_syev_select(int n, const double* A, double* w, int k, const int* sel, double* Z);
/* trl_hip_dense_syev_select for the projected matrix of the thick-restart
 * block Lanczos (block size b): A[0:p, 0:p] diagonal (the kept Ritz values),
 * those rows coupled only to the b coordinates p..p+b-1 (the arrow), and
 * A[p:, p:] of half-bandwidth b (block tridiagonal with upper-triangular
 * sub-diagonal blocks).  *p_out receives p (0 before the first restart), or
 * -1 with TRL_HIP_EINVAL when A lacks that structure.  The Ritz part is
 * reduced to band form by Householder QL steps, the band to tridiagonal form
 * by Householder bulge chasing: O(n^2 (b + k)) work in place
 * of the O(n^3) Householder tridiagonalisation.  The step it accelerates is
 * the dense eigensolve of impl/openmp/evp_base.hh:79-154; the reference has
 * no separate structured solver. */
int trl_hip_dense_syev_arrow_band(int n, int b, const double* A, double* w, int k, const int* sel, double* Z, int* p_out);

/* ---- Matrix Market coordinate files (host) -------------------------------
 * Replaces load_matrix_market of examples/sycl/csrevp.hh:26-124: '%' comment
 * lines, a "symmetric" banner mirrors the off-diagonal entries, 1-based
 * indices.  The CSR is canonical (columns ascending within a row, duplicate
 * entries summed); real, integer and pattern (value 1) fields, general or
 * symmetric.  trl_hip_mm_read returns TRL_HIP_EIO when the file cannot be
 * opened ("Error: cannot open file <path>", the reference's message) or is
 * malformed, unsupported or truncated; trl_hip_mm_last_error gives the
 * calling thread's message.  trl_hip_mm_copy fills caller-owned host arrays
 * of rows+1, nnz, nnz entries. */
typedef struct trl_hip_mm trl_hip_mm;
int trl_hip_mm_read(const char* path, trl_hip_mm** out);
int trl_hip_mm_info(const trl_hip_mm* m, size_t* rows, size_t* cols, size_t* nnz, int* symmetric);
int trl_hip_mm_copy(const trl_hip_mm* m, int32_t* rowptr, int32_t* col, double* val);
void trl_hip_mm_free(trl_hip_mm* m);
const char* trl_hip_mm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TRL_HIP_SOLVER_H */
