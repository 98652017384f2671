#pragma once
// Eigenproblem base of the HIP backend: the inner product, Cholesky-QR, the
// Rayleigh-Ritz step and the factories the Eigenproblem concept asks for
// (include/trl/concepts.hh:157-188), following impl/openmp/evp_base.hh:15-182.
//
// Differences from the OpenMP base, all deliberate:
//  - Scalar (eigenvalues, coefficients, tolerance) is double for both panel
//    precisions; BlockMultivector::Scalar is the panel type.
//  - orthonormalize runs Gram, a FP64 Cholesky and the scale on the device;
//    a breakdown is raised (same exception and message) at the next host
//    synchronisation instead of inside the call.
//  - the eigenvalue order is selectable: |lambda| descending (the OpenMP
//    default, evp_base.hh:103), ascending or descending algebraic (the SYCL
//    example's EigenvalueOrder, examples/sycl/evp_base.hh:24,110-117).
//  - an optional diagonal mass M turns dot into the M-inner product
//    (examples/openmp/csrevp.hh:89-97 with a diagonal B).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <memory>
#include <numeric>
#include <span>
#include <stdexcept>
#include <vector>

#include "trl_amd/rayleigh_ritz.hh"

#include "blockmatrix.hh"
#include "context.hh"
#include "multivector.hh"

namespace trl::hip {

enum class EigenvalueOrder { LargestMagnitude = 0, Ascending = 1, Descending = 2 };

template <class PanelScalar, unsigned int block_size>
class EVPBase {
public:
  using Scalar = double;
  using PanelType = PanelScalar;
  constexpr static auto blocksize = block_size;

  using BlockMultivector = ::trl::hip::BlockMultivector<PanelScalar, blocksize>;
  using BlockView = typename BlockMultivector::BlockView;
  using BlockMatrix = typename BlockMultivector::BlockMatrix;
  using BlockMatrixBlockView = typename BlockMatrix::BlockView;
  static constexpr trl_hip_dtype dtype = dtype_of<PanelScalar>();

  EVPBase(const Context& ctx, std::size_t n) : ctx_(ctx), N(n)
  {
    int g = 0;
    if (trl_hip_ctx_get_option(ctx_.get(), TRL_HIP_OPT_GRAPHS, &g) == TRL_HIP_OK) graphs_ = g != 0;
  }
  virtual ~EVPBase() = default;

  virtual void apply(BlockView X, BlockView Y) = 0;

  // R = V^T M W (evp_base.hh:37)
  virtual void dot(BlockView V, BlockView W, BlockMatrixBlockView R)
  {
    throw_on(trl_hip_gram(ctx_.get(), dtype, N, blocksize, 1, V.data_, W.data_, mass_, R.data_), "dot");
  }

  // V <- V R^{-1}, V_old = V_new R, R upper triangular (evp_base.hh:39-71)
  virtual void orthonormalize(BlockView V, BlockMatrixBlockView R)
  {
    throw_on(trl_hip_cholqr(ctx_.get(), dtype, N, blocksize, V.data_, mass_, R.data_, cholqr_passes_), "orthonormalize");
  }

  std::size_t size() const { return N; }

  BlockMultivector create_multivector(std::size_t rows, std::size_t cols) const { return BlockMultivector(ctx_, rows, cols); }

  BlockMatrix create_blockmatrix(std::size_t block_rows, std::size_t block_cols) const { return BlockMatrix(ctx_, block_rows, block_cols); }

  // Rayleigh-Ritz on the projected matrix and the converged count
  // (evp_base.hh:79-154): eigen-decompose T in FP64 on the host, order the
  // pairs, keep the eigenvectors on the device, and count the first nev
  // pairs with ||beta y_last|| / max(|theta|, eps) < tol.  All Ritz values
  // are computed, the Ritz vectors only for the first nev/b + 1 blocks (the
  // ones the driver reads) or set_ritz_blocks(), if more; the other columns
  // of the eigenvector matrix are zero.
  virtual std::size_t solve_small_dense(const BlockMatrix& T, BlockMatrixBlockView beta, std::size_t nev)
  {
    ctx_.sync(); // raises a pending Cholesky breakdown
    const auto t0 = std::chrono::steady_clock::now();
    const std::size_t nb = T.block_rows();
    const std::vector<double> tb = T.to_host();
    const std::vector<double> bh = beta.to_host();
    // the driver uses the first k_r = nev/b + 1 blocks of Ritz vectors
    // (restart rotation, arrow, final vectors: lanczos.hh:105-143); a caller
    // may ask for more (set_ritz_blocks)
    const std::size_t blocks = std::max<std::size_t>(nev / blocksize + 1, ritz_blocks_);
    trl_amd::rayleigh_ritz(blocksize, nb, T.block_cols(), tb.data(), bh.data(), nev, static_cast<int>(order_), tolerance_, rr_, dense_,
                           blocks * blocksize, !dense_rr_);
    ++rr_steps_;
    if (rr_.structured >= 0) ++rr_structured_;
    eigenvalues = rr_.theta;
    evec_host_ = rr_.Y;
    last_residuals_ = rr_.residual;
    if (!eigenvectors || eigenvectors->block_rows() != nb) eigenvectors = std::make_unique<BlockMatrix>(ctx_, nb, T.block_cols());
    eigenvectors->from_host(evec_host_);
    const std::size_t n_converged = rr_.nconv;
    host_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return n_converged;
  }

  const std::vector<double>& get_current_eigenvalues() const { return eigenvalues; }

  const BlockMatrix& get_current_eigenvectors() const
  {
    if (!eigenvectors) throw std::logic_error("no Rayleigh-Ritz solve yet");
    return *eigenvectors;
  }

  // The eigenvectors of the last Rayleigh-Ritz step, host copy.
  const std::vector<double>& get_current_eigenvectors_host() const { return evec_host_; }

  std::span<double, blocksize> get_eigenvalues_block(std::size_t block) { return std::span<double, blocksize>(eigenvalues.data() + block * blocksize, blocksize); }

  void set_tolerance(double tol) { tolerance_ = tol; }
  double get_tolerance() const { return tolerance_; }

  void set_order(EigenvalueOrder o) { order_ = o; }
  EigenvalueOrder get_order() const { return order_; }

  // 1: Cholesky-QR as in the reference; 2: CholQR2 (R = R2 R1)
  void set_cholqr_passes(int p) { cholqr_passes_ = p; }
  // At least this many leading blocks of Ritz vectors are computed by every
  // Rayleigh-Ritz step (default: the nev/b + 1 the driver uses).
  void set_ritz_blocks(std::size_t blocks) { ritz_blocks_ = blocks; }
  // true: the dense Householder eigensolver for every Rayleigh-Ritz step
  // (the structured one is used otherwise whenever T has the restart form)
  void set_dense_rayleigh_ritz(bool on) { dense_rr_ = on; }
  // Rayleigh-Ritz steps so far, and of them those by the structured solver
  unsigned rr_steps() const { return rr_steps_; }
  unsigned rr_structured() const { return rr_structured_; }
  int get_cholqr_passes() const { return cholqr_passes_; }

  // Optional diagonal inner-product weight M (device, panel type, n entries).
  void set_mass(const PanelScalar* m) { mass_ = m; }
  const PanelScalar* mass() const { return mass_; }

  const Context& context() const { return ctx_; }

  // Whether the driver may capture a cycle's block steps into a graph and
  // replay it (trl_hip_graph_*): true when apply only queues device work
  // (the CSR operators); an EVP whose apply reads back from the device says
  // false.  set_graphs turns the capture on or off (initially the context's
  // TRL_HIP_OPT_GRAPHS, off by default).
  virtual bool graph_capture_ok() const { return false; }
  void set_graphs(bool on) { graphs_ = on; }
  bool graphs() const { return graphs_; }
  // Changes whenever this EVP replaces a scratch buffer a captured step names.
  unsigned long long scratch_epoch() const { return coef_epoch_; }
  const std::vector<double>& last_residuals() const { return last_residuals_; }
  double host_seconds() const { return host_seconds_; }

  // ---- fused block products (used by BlockCGS2 and trl_amd::BlockLanczos) --

  // C = V[0:count]^T M W, C a device (count*b x b) FP64 array owned by this EVP.
  virtual double* gram_blocks(const BlockMultivector& V, unsigned count, const BlockView& W)
  {
    double* C = coef(std::size_t(count) * blocksize * blocksize);
    throw_on(trl_hip_gram(ctx_.get(), dtype, N, blocksize, int(count), V.data(), W.data_, mass_, C), "gram_blocks");
    return C;
  }

  // W orthogonalised against V[0:count] in the M-inner product by block
  // classical Gram-Schmidt: two passes (eta <= 0), or a second pass only where
  // the first one cancelled (the DGKS test with threshold eta), decided on the
  // device.
  virtual void reorth_blocks(const BlockMultivector& V, unsigned count, const BlockView& W, double eta)
  {
    double* C = coef(std::size_t(2 * count + 1) * blocksize * blocksize);
    throw_on(trl_hip_reorth(ctx_.get(), dtype, N, blocksize, int(count), V.data(), W.data_, mass_, C, eta), "reorth_blocks");
  }

  // W -= V[0:count] C
  void update_blocks(const BlockMultivector& V, unsigned count, const double* C, const BlockView& W)
  {
    throw_on(trl_hip_update(ctx_.get(), dtype, N, blocksize, int(count), V.data(), C, W.data_), "update_blocks");
  }

  // Out[0:kc] = V[0:m] Y[:, 0:kc] (lanczos.hh:105-114 as one contraction)
  void rotate(const BlockMultivector& V, unsigned m, const BlockMatrix& Y, unsigned kc, BlockMultivector& Out)
  {
    throw_on(trl_hip_rotate(ctx_.get(), dtype, N, blocksize, int(m), V.data(), Y.data(), int(Y.block_cols()), int(kc), Out.data()), "rotate");
  }

protected:
  double* coef(std::size_t count)
  {
    if (coef_.size() < count) {
      coef_ = DeviceBuffer<double>(ctx_, std::max<std::size_t>(count, 4096));
      ++coef_epoch_;
    }
    return coef_.data();
  }

  Context ctx_;
  std::size_t N;
  const PanelScalar* mass_ = nullptr;
  int cholqr_passes_ = std::is_same_v<PanelScalar, float> ? 2 : 1;
  std::size_t ritz_blocks_ = 0;
  bool dense_rr_ = false;
  bool graphs_ = false;
  unsigned long long coef_epoch_ = 0;
  unsigned rr_steps_ = 0, rr_structured_ = 0;

  std::vector<double> eigenvalues;
  std::unique_ptr<BlockMatrix> eigenvectors;
  std::vector<double> evec_host_;
  std::vector<double> dense_;
  std::vector<double> last_residuals_;
  trl_amd::RitzResult rr_;
  double tolerance_ = 1e-8;
  EigenvalueOrder order_ = EigenvalueOrder::LargestMagnitude;
  double host_seconds_ = 0.0;
  DeviceBuffer<double> coef_;
};

} // namespace trl::hip
