/*
 * trl_hip.h — C-ABI of libtrl_hip.so, the MI355X (gfx950) kernel library
 * behind the trl::hip block-Lanczos backend.
 *
 * Every entry point replaces one operation of trl's backend concepts
 * (include/trl/concepts.hh in the reference) or of its OpenMP implementation
 * (the include/trl/impl/openmp headers); the reference location is cited beside each
 * declaration.  Plain pointers and sizes only: no torch or C++ types cross.
 *
 * Conventions
 *  - A "panel" is an n x b block of a multivector, row-major: element (r, c)
 *    at offset r*b + c (impl/openmp/multivector.hh:88).  Block j of a
 *    multivector starts j*n*b elements after block 0.
 *  - Small b x b blocks ("coefficient blocks", the MatrixBlockView of the
 *    concepts) are row-major FP64 in device memory for both panel types:
 *    the FP32 form keeps every coefficient in FP64.
 *  - dtype selects the panel scalar (FP32 or FP64).  CSR values, diagonal
 *    scalings and mass diagonals have the panel dtype; indices are int32.
 *  - b must be one of 1, 2, 4, 6, 8, 16.
 *  - All calls are ordered on the context's single stream and return as soon
 *    as the work is queued, except those documented as synchronous.
 *  - Return value: 0 = success, >0 = hipError_t, <0 = TRL_HIP_E* below.
 *    A Cholesky breakdown inside a queued orthonormalization cannot be
 *    returned by the call that queued it; it is recorded in a sticky device
 *    flag and reported as TRL_HIP_ECHOL by the next trl_hip_ctx_check().
 *  - One context per host thread; a context is not thread-safe.
 */
#ifndef TRL_HIP_H
#define TRL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRL_HIP_OK 0
#define TRL_HIP_EINVAL (-1) /* bad argument (std::invalid_argument in the reference) */
#define TRL_HIP_ECHOL (-2)  /* Cholesky breakdown (evp_base.hh:55 runtime_error)     */
#define TRL_HIP_ENOMEM (-3) /* device allocation failed                               */
#define TRL_HIP_EEIG (-4)   /* dense eigensolver did not converge (evp_base.hh:96)   */
#define TRL_HIP_ERANGE (-5) /* block index out of range (multivector.hh:87)           */
#define TRL_HIP_EBARRIER (-6) /* a batched step's device barrier timed out             */
#define TRL_HIP_EIO (-7)      /* a file cannot be opened or parsed                      */
#define TRL_HIP_ENOCONV (-8)  /* an inner iterative solve missed its tolerance          */
#define TRL_HIP_EINDEF (-9)   /* an operator that must be SPD gave p^T K p <= 0         */
#define TRL_HIP_ESTALE (-10)  /* a graph names scratch memory since replaced            */

typedef enum { TRL_HIP_F32 = 0, TRL_HIP_F64 = 1 } trl_hip_dtype;

#define TRL_HIP_PROF_FAMILIES 8

typedef struct trl_hip_ctx trl_hip_ctx;

/* ---- context, memory ---------------------------------------------------- */

/* Creates a context on `device` with one in-order stream (the in-order SYCL
 * queue of impl/sycl/blockview.hh:19-22 is the reference's device analogue). */
int trl_hip_ctx_create(int device, trl_hip_ctx** out);
int trl_hip_ctx_destroy(trl_hip_ctx* ctx);
/* Waits for all queued work. */
int trl_hip_ctx_sync(trl_hip_ctx* ctx);
/* Waits, then reads and clears the sticky Cholesky flag.  Returns
 * TRL_HIP_ECHOL (and *pivot = failing pivot index, 1-based) if any
 * orthonormalization since the last check broke down. */
int trl_hip_ctx_check(trl_hip_ctx* ctx, int* pivot);
/* The context's hipStream_t, for callers that interleave their own work. */
void* trl_hip_ctx_stream(trl_hip_ctx* ctx);
/* Number of compute units of the context's device. */
int trl_hip_ctx_num_cus(trl_hip_ctx* ctx);
/* Context options (no reference counterpart; for tests of paths that the
 * problem sizes at hand would not take):
 *   TRL_HIP_OPT_OFFSETS64, value 1: the operator kernels use their 64-bit
 *   element offsets, which they otherwise take only when a panel or an ELL
 *   array exceeds 4 GiB, and the Ritz rotation its 64-bit-address form,
 *   otherwise taken only when its buffers exceed 2 GiB; value 0 restores the
 *   size-based choice. */
#define TRL_HIP_OPT_OFFSETS64 1
/*   TRL_HIP_OPT_GRAPHS, value 1: eigenproblems created on the context let
 *   the driver capture each restart cycle's block steps into a graph once and
 *   replay it (the same kernels in the same order); 0 (default): the steps
 *   are queued directly.  Measured on the 128^2 Laplacian (~15 kernels of a
 *   few microseconds per block step) the replay is no faster: the device,
 *   not the host, paces such steps. */
#define TRL_HIP_OPT_GRAPHS 2
/*   TRL_HIP_OPT_ALLOC_CACHE, value 1 (default): trl_hip_free keeps a block
 *   of trl_hip_malloc for reuse by a later allocation of the same size on the
 *   context (zero-filled again; reuse is ordered on the context stream, so
 *   the free neither synchronises nor calls hipFree); the blocks, at most 64
 *   GiB, are released when an allocation fails, at trl_hip_ctx_destroy, or
 *   when the option is set to 0. */
#define TRL_HIP_OPT_ALLOC_CACHE 3
int trl_hip_ctx_set_option(trl_hip_ctx* ctx, int option, int value);
int trl_hip_ctx_get_option(trl_hip_ctx* ctx, int option, int* value);
const char* trl_hip_strerror(int code);

/* Zero-initialised device allocation (multivector.hh:44-48 zero-fills).
 * trl_hip_free of a block from trl_hip_malloc returns it to the context's
 * cache (TRL_HIP_OPT_ALLOC_CACHE); other pointers are freed after the
 * context stream has finished. */
int trl_hip_malloc(trl_hip_ctx* ctx, size_t bytes, void** out);
int trl_hip_free(trl_hip_ctx* ctx, void* p);
/* Synchronous with respect to the host: return after the copy. */
int trl_hip_memcpy_h2d(trl_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int trl_hip_memcpy_d2h(trl_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Queued device-to-device copy: BlockView::copy_from (blockview.hh:59-63). */
int trl_hip_memcpy_d2d(trl_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Queued zero fill by a kernel (no memset node, so it is safe inside a
 * captured graph): BlockView::set_zero (blockview.hh:31-36). */
int trl_hip_fill_zero(trl_hip_ctx* ctx, void* dst, size_t bytes);
/* Element conversion between panel dtypes (count elements). */
int trl_hip_convert(trl_hip_ctx* ctx, trl_hip_dtype src_t, const void* src, trl_hip_dtype dst_t, void* dst, size_t count);

/* ---- panel operations (BlockVectorView, concepts.hh:85-106) ------------- */

/* X -= Y over n*b elements: BlockView::operator-= (blockview.hh:69-74). */
int trl_hip_panel_sub(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const void* Y);

#define TRL_HIP_MULT 0           /* Y  = X * W   : BlockView::mult (blockview.hh:104-119)            */
#define TRL_HIP_MULT_ADD 1       /* Y += X * W   : BlockView::mult_add (blockview.hh:81-97)          */
#define TRL_HIP_MULT_TRANSPOSE 2 /* Y  = X * W^T : BlockView::mult_transpose (blockview.hh:126-141)  */
#define TRL_HIP_SUBTRACT_PRODUCT 3 /* X -= Y * W : BlockView::subtract_product (blockview.hh:189-201) */
#define TRL_HIP_SUBTRACT_PRODUCT_TRANSPOSE 4 /* X -= Y * W^T : mult_transpose then -= (lanczos.hh:255-258) in one pass */
/* W: b x b FP64 device block. */
int trl_hip_panel_mult(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const double* W, void* Y, int mode);

/* ---- block products (the Gram / update contractions) -------------------- */

/* C = V[0:k]^T diag(m) W, C is (k*b) x b FP64 on the device, block j of C
 * (rows j*b..j*b+b-1) = V_j^T diag(m) W.  mdiag may be NULL (m = 1).
 * k = 1 is BlockView::dot / EVPBase::dot (blockview.hh:152-182,
 * evp_base.hh:37); mdiag != NULL is the B-inner product of
 * CSRGeneralizedEVP::dot (examples/openmp/csrevp.hh:89-97) with a diagonal B.
 * k > 1 is the fused form of the ModifiedGS dot loop
 * (reorthogonalization.hh:14-21). */
int trl_hip_gram(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const void* W, const void* mdiag, double* C);

/* trl_hip_gram without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and result as trl_hip_gram, with the per-workgroup partial sums in
 * a fixed 8 MiB device array of the library (one per device, held until the
 * process ends, its address resolved at trl_hip_ctx_create) instead of the
 * context's growable buffer.  k*b*b must not exceed
 * TRL_HIP_NOALLOC_MAX_COEFS (TRL_HIP_EINVAL otherwise).  Every call on one
 * device shares that array: calls that overlap on one device (several
 * contexts or host threads on the same GPU) race. */
#define TRL_HIP_NOALLOC_MAX_COEFS 4096
int trl_hip_gram_noalloc(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const void* W, const void* mdiag, double* C);

/* W -= sum_{j<k} V_j * C_j with C as returned by trl_hip_gram: the fused
 * form of the subtract_product loop of ModifiedGS (reorthogonalization.hh:19). */
int trl_hip_update(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const double* C, void* W);

/* Cholesky-QR of one panel: W <- W R^{-1}, R upper triangular with positive
 * diagonal and R^T R = W^T diag(m) W, so that W_old = W_new R
 * (concepts.hh:153-155; EVPBase::orthonormalize, evp_base.hh:39-71).
 * The b x b Cholesky runs in FP64 on the device.  passes = 1 is the
 * reference's single CholQR; passes = 2 repeats it on the result and returns
 * R = R2 R1 (CholQR2).  R: b x b FP64 device block.  On breakdown W is left
 * unchanged and the sticky flag is set (see trl_hip_ctx_check). */
int trl_hip_cholqr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, const void* mdiag, double* R, int passes);

/* Thick-restart Ritz rotation (lanczos.hh:105-114 as one contraction):
 * Out_j = sum_{i<m} V_i * Y(i, j) for j < kc, where Y is a FP64 device block
 * matrix in BlockMatrix layout (block (i, j) at Y + (i*y_block_cols + j)*b*b,
 * impl/openmp/blockmatrix.hh:52-58).  Out holds kc panels and must not
 * overlap V. */
int trl_hip_rotate(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int m, const void* V, const double* Y, int y_block_cols, int kc, void* Out);

/* Block classical Gram-Schmidt of W against V[0:k] in the M-inner product
 * (the Reorth hook, reorthogonalization.hh:29-32; the reference's single
 * block-MGS pass is reorthogonalization.hh:12-22): C = V^T M W,
 * W -= V C, and a second pass with fresh coefficients
 *   eta <= 0:     always (CGS2);
 *   0 < eta < 1:  only if some column's norm after the first pass is below
 *                 eta times its norm before (the DGKS test, Daniel, Gragg,
 *                 Kaufman & Stewart 1976), decided on the device: the
 *                 second pass's kernels return at once when it is not needed.
 * When W is the block right after V[0:k] (W == V + k*n*b) the first pass
 * also yields W^T M W in the same sweep.  C: device scratch of
 * (2k+1)*b*b doubles. */
int trl_hip_reorth(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, void* W, const void* mdiag, double* C, double eta);

/* CholQR in the inner product of a sparse SPD matrix B: the generalized
 * problem's orthonormalize (EVPBase::orthonormalize, evp_base.hh:39-71, with
 * the B-dot of examples/openmp/csrevp.hh:89-97), given BW = B W (n x b,
 * device, not overlapping W).  Per pass G = W^T BW, G = R^T R (FP64,
 * breakdown as trl_hip_cholqr), W <- W R^{-1} and BW <- BW R^{-1}; with
 * passes = 2 the returned R is R2 R1. */
int trl_hip_cholqr_bprod(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, void* BW, double* R, int passes);

/* ---- captured graphs -------------------------------------------------------- */

/* Stream capture of queued work into a graph that the context replays with
 * one launch (no reference counterpart: the host cost of a Lanczos cycle's
 * ~20 launches per block step, or 2(i+1) + 8 per step with ModifiedGS,
 * replaced by one graph launch).  trl_hip_graph_begin starts capturing the
 * context stream: the calls that follow queue nothing until
 * trl_hip_graph_end, which returns an executable graph, or an error and no
 * graph when a captured call needed the host (a synchronisation, a device
 * allocation, a read back) — the captured work then has not run.  Not in
 * profiling mode.  trl_hip_graph_launch queues one replay on the context
 * stream: the same kernels on the same device addresses and arguments;
 * TRL_HIP_ESTALE if the context has since replaced a scratch buffer the
 * graph names.  trl_hip_graph_destroy frees the graph. */
typedef struct trl_hip_graph trl_hip_graph;
int trl_hip_graph_begin(trl_hip_ctx* ctx);
int trl_hip_graph_end(trl_hip_ctx* ctx, trl_hip_graph** graph);
int trl_hip_graph_launch(trl_hip_ctx* ctx, const trl_hip_graph* graph);
int trl_hip_graph_destroy(trl_hip_graph* graph);

/* ---- operator ------------------------------------------------------------ */

/* A sparse operator as the kernels see it (plain data, caller-owned device
 * arrays).  format 0 = CSR (rowptr[n+1], col/val[nnz], sorted columns),
 * 1 = slot-major ELL (col/val[width*n], slot s of row r at s*n + r, padding
 * slots col = r, val = 0).  dscale (optional, panel dtype, n entries) scales
 * the rows of the product: dscale = 1/diag(M) applies M^{-1}K for a
 * generalized problem with a diagonal M.  variant selects the kernel:
 *   0 default, 1 one row per lane, -2 / -4 two / four lanes per row splitting
 *   the b columns, 2 / 4 / 8 / 16 lane groups splitting the nonzeros (CSR).
 * Indices and values are loaded in chunks of 8 nonzeros before their gathers.
 * plane / plane_block select the order in which rows are visited (the
 * product is the same): plane = 0 visits rows in index order; plane > 0
 * views row r as (r / plane, r % plane) and visits blocks of plane_block
 * consecutive rows of every plane in turn (block 0 of all planes, then
 * block 1, ...).  For a matrix whose far off-diagonals sit +-plane away (a
 * grid stencil), the rows of X that a block gathers are then reused while
 * they are cached instead of two planes later.  0 < plane_block <= plane. */
typedef struct {
  int format;
  const int32_t* rowptr;
  int width;
  const int32_t* col;
  const void* val;
  const void* dscale;
  int variant;
  int plane;
  int plane_block;
  /* The several-panel ELL step only (trl_hip_spmm_panels and its axpby and
   * forced forms on ELL rows of at most 8 slots; ignored elsewhere):
   *   share = 1: a gathered column that is a row of the same wave (the +-1
   *     neighbours of a stencil row, the diagonal, padding) is taken from the
   *     lane that loads that row as its own, by a cross-lane permute, instead
   *     of a load of its own;
   *   line > 0, tile_lines > 0 (with plane > 0): inside each plane block a
   *     workgroup visits tile_lines consecutive lines of `line` rows instead
   *     of consecutive rows, so that the rows gathered +-line away are mostly
   *     its own; used only where such tiles cover every plane block exactly
   *     (index order otherwise).  The product is the same. */
  int share;
  int line;
  int tile_lines;
} trl_hip_spmat;

/* Y = diag(dscale) A X: EVP::apply of CSREVP (examples/openmp/csrevp.hh:36-44;
 * SYCL kernel examples/sycl/csrevp.hh:159-187). */
int trl_hip_spmm(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, const void* X, void* Y);

/* Fused three-term recurrence step: Y = alpha diag(dscale) A X + beta X +
 * gamma Z (Z may be NULL, or equal to Y for an in-place update of the older
 * iterate; X must not overlap Y).  One step of a Chebyshev polynomial filter;
 * no reference counterpart. */
int trl_hip_spmm_axpby(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, const void* X, void* Y, const void* Z, double alpha,
                       double beta, double gamma);

/* The same two products on npanels panels at once (panel j of X, Y, Z at
 * j*n*b elements): each lane applies its matrix row to every panel in turn,
 * so the matrix entries are fetched once for all panels (ELL rows of at most
 * 8 slots keep them in registers).  For the Chebyshev filter of a block of
 * Ritz vectors. */
int trl_hip_spmm_panels(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, int npanels, const void* X, void* Y);
int trl_hip_spmm_axpby_panels(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, int npanels, const void* X, void* Y, const void* Z,
                              double alpha, double beta, double gamma);

/* The panel step with a forcing term: Y_j = alpha diag(dscale) A X_j +
 * beta X_j + gamma Z_j + F_j diag(g_j) for j < npanels, F laid out as X,
 * g device doubles, g[j*b + c] the coefficient of column c of panel j.  The
 * step of the residual-driven Chebyshev filter (E_i = p_i(A) x - p_i(theta)
 * x for a Ritz pair (theta, x), forced by the residual A x - theta x), which
 * lets an FP32 recurrence refine FP64 Ritz vectors.  No reference
 * counterpart. */
int trl_hip_spmm_axpby_forced(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, int npanels, const void* X, void* Y, const void* Z,
                              double alpha, double beta, double gamma, const void* F, const double* g);

/* Residual panels of Ritz pairs for that filter: R_j = AX_j - X_j
 * diag(theta_j) computed in FP64 (AX, X: npanels FP64 panels; theta: device,
 * npanels*b) and stored in FP32 as R and, times s, as E. */
int trl_hip_residual_split(trl_hip_ctx* ctx, size_t n, int b, int npanels, const double* AX, const double* X, const double* theta, double s, float* R,
                           float* E);

/* X_j = X_j diag(h_j) + E_j for j < npanels (X FP64, E FP32, h device
 * doubles, npanels*b): the filtered Ritz vectors p(A) x = p(theta) x + E. */
int trl_hip_panel_combine(trl_hip_ctx* ctx, size_t n, int b, int npanels, double* X, const double* h, const float* E);

/* CSR convenience forms of the two calls above (variant as in trl_hip_spmat). */
int trl_hip_spmm_csr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const int32_t* rowptr, const int32_t* col, const void* val, const void* X, void* Y,
                     const void* dscale, int variant);
int trl_hip_spmm_csr_axpby(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const int32_t* rowptr, const int32_t* col, const void* val, const void* dscale,
                           const void* X, void* Y, const void* Z, double alpha, double beta, double gamma);
/* ELL convenience form (default variant). */
int trl_hip_spmm_ell(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int width, const int32_t* ell_col, const void* ell_val, const void* X, void* Y,
                     const void* dscale);

/* Synchronous: the longest row of a CSR matrix (the ELL width). */
int trl_hip_csr_max_row(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, int* width);
/* Queued: CSR -> slot-major ELL of the given width (>= the longest row);
 * ell_col/ell_val hold width*n entries. */
int trl_hip_csr_to_ell(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, const void* val, int width, int32_t* ell_col,
                       void* ell_val);
/* Synchronous autotuner of the operator kernel for one matrix and block
 * size: times kernel variants of `csr` and, if ell is not NULL, of `ell`
 * (same matrix in ELL form), and blocked row orders (plane = the matrix
 * bandwidth) for the fastest of them, with HIP events on scratch panels;
 * a format whose first variant is far behind is not searched further.
 * Writes the fastest descriptor to *best (dscale copied from csr) and its
 * mean time per call to *best_us (may be NULL). */
int trl_hip_spmm_tune(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* csr, const trl_hip_spmat* ell, trl_hip_spmat* best,
                      double* best_us);
/* The same for the fused axpby form (trl_hip_spmm_axpby with Z = Y, the
 * Chebyshev filter's in-place step), whose best variant can differ (it reads
 * and writes more panel bytes per row). */
int trl_hip_spmm_tune_axpby(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* csr, const trl_hip_spmat* ell, trl_hip_spmat* best,
                            double* best_us);
/* The same for the axpby form on npanels panels at once
 * (trl_hip_spmm_axpby_panels with Z = Y); *best_us is per launch. */
int trl_hip_spmm_tune_panels(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int npanels, const trl_hip_spmat* csr, const trl_hip_spmat* ell,
                             trl_hip_spmat* best, double* best_us);

/* ---- shift-invert inner solve ---------------------------------------------- */

/* X = K^{-1} rhs for b right-hand sides (n x b panels, FP64) with K = A -
 * sigma B symmetric positive definite: the solve of the reference's
 * shift-invert apply (examples/openmp/csrevp.hh:76-84, Eigen SparseLU there)
 * by the Jacobi-preconditioned conjugate gradient method, one recurrence per
 * column, on the device (K any trl_hip_spmat form, dinv = 1/diag(K), n
 * doubles).  Stops when every column's recurrence residual satisfies
 * |r| <= tol |rhs|, or after maxit iterations (TRL_HIP_ENOCONV).
 * TRL_HIP_EINDEF when a search direction gives p^T K p <= 0.  work: device
 * scratch of trl_hip_pcg_workspace(n, b) bytes.  Synchronous (the host reads
 * the residuals every 8 iterations); *iters and *relres (max over columns of
 * |r| / |rhs|) may be NULL.  No reference counterpart for the iteration
 * itself (the reference factorises). */
size_t trl_hip_pcg_workspace(size_t n, int b);
int trl_hip_pcg(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* K, const void* dinv, const void* rhs, void* X, double tol,
                int maxit, void* work, int* iters, double* relres);

/* ---- small b x b blocks (MatrixBlockViewConcept, concepts.hh:26-39) ----- */

/* dst = src (transpose = 0) or src^T (transpose = 1):
 * BlockMatrixBlockView::copy_from / copy_from_transpose (matrixblockview.hh:30-43). */
int trl_hip_small_copy(trl_hip_ctx* ctx, int b, const double* src, double* dst, int transpose);
/* dst = 0: BlockMatrixBlockView::set_zero (matrixblockview.hh:49). */
int trl_hip_small_zero(trl_hip_ctx* ctx, int b, double* dst);
/* diag(dst) = host_values[0..b): set_diagonal (matrixblockview.hh:71-74). */
int trl_hip_small_set_diagonal(trl_hip_ctx* ctx, int b, double* dst, const double* host_values);
/* C = A B: BlockMatrixBlockView::mult (matrixblockview.hh:55-64). */
int trl_hip_small_mult(trl_hip_ctx* ctx, int b, const double* A, const double* B, double* C);

/* ---- eigenvalue adjoint (no reference counterpart) ------------------------ */

/* SDDMM on the pattern of A: for every stored entry p = (i, j) of the CSR
 * pattern, G[p] = sum_q w_q X[i][q] X[j][q], the gradient of the loss
 * sum_q w_q lambda_q with respect to A_ij for unit (or M-unit) eigenvectors
 * (d lambda_q / d A_ij = x_iq x_jq).  X holds nvec = nblocks*b eigenvectors as
 * nblocks row-major n x b panels (block j at X + j*n*b), w is a device FP64
 * array of nvec weights, G a device FP64 array of nnz entries.  If gm is not
 * NULL it receives, per row i, the mass-diagonal gradient of a generalized
 * problem, gm[i] = -sum_q w_q lambda_q X[i][q]^2 (lambda: device FP64, nvec). */
int trl_hip_sddmm_adjoint(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, int b, int nblocks, const void* X,
                          const double* w, const double* lambda, double* G, double* gm);

/* ---- checks -------------------------------------------------------------- */

/* Synchronous.  For the b columns c of panel X and of Y:
 * out[c] = || Y[:,c] - lambda[c] * diag(m) X[:,c] ||_2^2 and
 * out[b + c] = || Y[:,c] ||_2^2, accumulated in FP64.  lambda: host array of b
 * values (may be NULL: lambda = 0).  mdiag may be NULL.  Used for the true
 * residual of Ritz pairs (Y = A X). */
int trl_hip_residual_sq(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const void* X, const void* Y, const void* mdiag, const double* lambda, double* out);

/* ---- batched problems (no reference counterpart) --------------------------
 *
 * Many independent standard eigenproblems A_p x = lambda x (FP64, the same
 * block size b and basis size ncv) advanced in lock-step by the thick-restart
 * block Lanczos recurrence.  The problems' rows are concatenated; basis block
 * q of all problems is one panel.  One Lanczos step of every active problem
 * is one persistent kernel launch (rows of the new block kept in registers
 * across the three-term recurrence, two classical Gram-Schmidt passes and
 * Cholesky-QR, workgroups of a problem synchronised by device barriers); the
 * steps of a restart cycle are replayed from a captured hipGraph.  The host
 * driver (trl_hip_batch_solve_csr in trl_hip_solver.h) runs the
 * Rayleigh-Ritz steps.  b in {1, 2, 4, 8}.  All calls are synchronous. */
typedef struct trl_hip_batch trl_hip_batch;

/* Host CSR arrays of nprob problems (n[p] rows, sorted columns); ncv a
 * multiple of b.  Copies them into the concatenated device layout (CSR and
 * ELL) and picks the default kernel variant. */
int trl_hip_batch_create(trl_hip_ctx* ctx, int nprob, const size_t* n, const int32_t* const* rowptr, const int32_t* const* col, const double* const* val,
                         int b, int ncv, trl_hip_batch** out);
int trl_hip_batch_destroy(trl_hip_batch* batch);
/* Kernel variant: format 0 CSR / 1 ELL, rows_per_thread (tile = 256 x that
 * many rows); tiles cover the active problems only; TRL_HIP_ERANGE if the
 * grid would not fit on the device at once. */
int trl_hip_batch_set_variant(trl_hip_batch* batch, int format, int rows_per_thread);
/* Autotuner: times one full-width step of the active problems for every
 * variant that fits and keeps the fastest (uses block m of the basis and the
 * coefficient arrays as scratch: call it before an extend that starts at
 * block 0 or with the quasi-Lanczos step). */
int trl_hip_batch_tune(trl_hip_batch* batch, double* best_us);
int trl_hip_batch_info(const trl_hip_batch* batch, int* ntiles, int* format, int* rows_per_thread, double* tuned_us);
/* cholqr_passes 1 (as the reference's orthonormalize) or 2; use_graph 0
 * launches the steps directly instead of replaying a captured graph. */
int trl_hip_batch_set_options(trl_hip_batch* batch, int cholqr_passes, int use_graph);
/* adaptive = 0: two classical Gram-Schmidt passes in every step (the
 * default); 1: the second pass only for the problems whose first pass
 * cancelled a column below 1/sqrt(2) of its norm (the DGKS test of
 * trl_hip_reorth with eta = 1/sqrt(2)), decided on the device. */
int trl_hip_batch_set_reorth(trl_hip_batch* batch, int adaptive);
/* Host <-> device copy of basis block q (n[p] x b) of problem p. */
int trl_hip_batch_set_block(trl_hip_batch* batch, int p, int q, const double* host);
int trl_hip_batch_get_block(trl_hip_batch* batch, int p, int q, double* host);
/* active[p] != 0: problem p takes part in the following steps/rotations. */
int trl_hip_batch_set_active(trl_hip_batch* batch, const int* active);
/* Steps k..m_end-1 of every active problem (lanczos.hh:227-287): k == 0
 * orthonormalises block 0 first; quasi = 1 makes step k the quasi-Lanczos
 * step after a restart (lanczos.hh:150-178) with arrow[p] = T_{k,k-1} (b x b,
 * nprob blocks).  Returns alpha (T_ii) and beta (the R of orthonormalising
 * V_{i+1}) as nprob x m blocks (step i at [p*m + i]) and the sticky Cholesky
 * breakdown pivots per problem (0 = none). */
int trl_hip_batch_extend(trl_hip_batch* batch, int k, int m_end, int quasi, const double* arrow, double* alpha, double* beta, int* pflag);
/* Thick restart of every active problem: block j < kr becomes
 * sum_{q<m} V_q Y_p(q, j), block kr becomes V_m (lanczos.hh:105-118).  Y:
 * nprob row-major (m b) x (kr b) matrices. */
int trl_hip_batch_rotate(trl_hip_batch* batch, const double* Y, int kr);
/* Copies blocks 0..nblk-1 of problem p out of the basis into the batch's
 * result array (done when p finishes: later steps and restarts only cover
 * active problems); trl_hip_batch_get_block then reads from there. */
int trl_hip_batch_save_blocks(trl_hip_batch* batch, int p, int nblk);
/* ||A x - lambda x|| / ||A x|| of the pairs (block q column c of every
 * problem, lambda[(p*nblk + q)*b + c]) into out (same indexing). */
int trl_hip_batch_residuals(trl_hip_batch* batch, int nblk, const double* lambda, double* out);
/* Clears the breakdown pivots and the barrier error flag. */
int trl_hip_batch_reset_flags(trl_hip_batch* batch);
/* Sets the barrier error flag as a timed-out device barrier does (for tests
 * of the error path): the next trl_hip_batch_extend then returns
 * TRL_HIP_EBARRIER, its step kernels returning at once, until
 * trl_hip_batch_reset_flags. */
int trl_hip_batch_raise_error(trl_hip_batch* batch);

/* ---- profiling (no reference counterpart) --------------------------------- */

/* Kernel families timed in profiling mode.  Algorithmic bytes per call
 * (P = n*b*sizeof(scalar)): SPMM nnz*(sizeof(val)+4) + (n+1)*4 + 2P
 * (+ n*sizeof(scalar) with a row scaling); GRAM (k+1)P (+ mass); UPDATE
 * (k+2)P; CHOLQR 3P per pass (+ mass); ROTATE (m+kc)P; PANEL 2P or 3P. */
#define TRL_HIP_PROF_SPMM 0
#define TRL_HIP_PROF_GRAM 1
#define TRL_HIP_PROF_UPDATE 2
#define TRL_HIP_PROF_CHOLQR 3
#define TRL_HIP_PROF_ROTATE 4
#define TRL_HIP_PROF_PANEL 5
#define TRL_HIP_PROF_SMALL 6
#define TRL_HIP_PROF_OTHER 7

/* on = 1 clears the counters and starts recording a HIP event pair on the
 * context stream around every call of the families above; on = 0 stops. */
int trl_hip_prof_enable(trl_hip_ctx* ctx, int on);
/* Synchronous: folds finished calls in and returns the totals of one family
 * since the last enable: calls, seconds between the event pairs, bytes. */
int trl_hip_prof_read(trl_hip_ctx* ctx, int family, long long* calls, double* seconds, double* bytes);
/* The same totals restricted to calls on panels of dtype t. */
int trl_hip_prof_read_dtype(trl_hip_ctx* ctx, int family, trl_hip_dtype t, long long* calls, double* seconds, double* bytes);
/* Profiling mode: trl_hip_reorth calls with eta > 0 since the last enable,
 * and how many of them took the second pass. */
int trl_hip_prof_reorth(trl_hip_ctx* ctx, long long* calls, long long* second_passes);
/* Region timer: begin records a HIP event on the context stream, end records
 * a second one, waits for it and returns the seconds between the two (the
 * device time of everything queued in between, without per-call events). */
int trl_hip_region_begin(trl_hip_ctx* ctx);
int trl_hip_region_end(trl_hip_ctx* ctx, double* seconds);

#ifdef __cplusplus
}
#endif

#endif /* TRL_HIP_H */
