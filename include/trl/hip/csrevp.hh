#pragma once
// Sparse eigenproblems on the HIP backend.
//
// CSREVP follows examples/openmp/csrevp.hh:19-48 (Y = A X for a symmetric
// CSR matrix); with a diagonal mass M it is the generalized problem
// K x = lambda M x in the form A = M^{-1} K, self-adjoint in the M-inner
// product (the B-inner product of examples/openmp/csrevp.hh:89-97 with a
// diagonal B; no factorisation is needed).  Arrays are device arrays of the
// panel type, either borrowed from the caller or uploaded and owned.

#include <algorithm>
#include <climits>
#include <cstdint>
#include <string>
#include <vector>

#include "evp_base.hh"
#include "sparse_operator.hh"

namespace trl::hip {

template <class PanelScalar, unsigned int block_size>
class CSREVP : public EVPBase<PanelScalar, block_size> {
public:
  using Base = EVPBase<PanelScalar, block_size>;
  using BlockView = typename Base::BlockView;

  // Borrows device arrays (rowptr[n+1], col[nnz], val[nnz]; mass and its
  // reciprocal optional, both or neither).
  CSREVP(const Context& ctx, std::size_t n, const int32_t* rowptr, const int32_t* col, const PanelScalar* val, const PanelScalar* mass = nullptr,
         const PanelScalar* mass_inv = nullptr)
      : Base(ctx, n), rowptr_(rowptr), col_(col), val_(val), minv_(mass_inv),
        op_(std::make_shared<SparseOperator<PanelScalar>>(ctx, n, rowptr, col, val, mass_inv))
  {
    if ((mass == nullptr) != (mass_inv == nullptr)) throw std::invalid_argument("CSREVP: give both the mass diagonal and its reciprocal, or neither");
    this->set_mass(mass);
  }

  // Uploads host CSR arrays (values and mass converted to the panel type).
  static std::unique_ptr<CSREVP> from_host(const Context& ctx, std::size_t n, const std::vector<int32_t>& rowptr, const std::vector<int32_t>& col,
                                           const std::vector<double>& val, const std::vector<double>* mass = nullptr)
  {
    if (rowptr.size() != n + 1 || col.size() != val.size() || std::size_t(rowptr.back()) != val.size())
      throw std::invalid_argument("CSREVP::from_host: inconsistent CSR arrays");
    auto owned = std::make_shared<Owned>(ctx, n, rowptr, col, val, mass);
    std::unique_ptr<CSREVP> e(new CSREVP(ctx, n, owned->rowptr.data(), owned->col.data(), owned->val.data(), mass ? owned->mass.data() : nullptr,
                                         mass ? owned->minv.data() : nullptr));
    e->owned_ = owned;
    return e;
  }

  void apply(BlockView X, BlockView Y) override { op_->apply(int(block_size), X.data_, Y.data_); }

  // apply only queues the operator kernel: a cycle's block steps can be captured
  bool graph_capture_ok() const override { return true; }

  // Fixed kernel variant of the CSR operator (trl_hip_spmat::variant).
  void set_lanes_per_row(int l) { op_ = std::make_shared<SparseOperator<PanelScalar>>(this->ctx_, this->N, rowptr_, col_, val_, minv_, l); }

  // Times the operator kernels (CSR variants and ELL) for this block size and
  // keeps the fastest; returns its time per apply in microseconds.
  double autotune(bool allow_ell = true) { return op_->tune(int(block_size), allow_ell); }

  // Shares an operator (for instance one tuned before) over the same matrix.
  void set_operator(std::shared_ptr<SparseOperator<PanelScalar>> op) { op_ = std::move(op); }
  const std::shared_ptr<SparseOperator<PanelScalar>>& op() const { return op_; }

  const int32_t* rowptr() const { return rowptr_; }
  const int32_t* col() const { return col_; }
  const PanelScalar* val() const { return val_; }
  const PanelScalar* mass_inv() const { return minv_; }

private:
  struct Owned {
    DeviceBuffer<int32_t> rowptr, col;
    DeviceBuffer<PanelScalar> val, mass, minv;
    Owned(const Context& ctx, std::size_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& c, const std::vector<double>& v,
          const std::vector<double>* m)
        : rowptr(ctx, rp.size()), col(ctx, std::max<std::size_t>(c.size(), 1)), val(ctx, std::max<std::size_t>(v.size(), 1))
    {
      throw_on(trl_hip_memcpy_h2d(ctx.get(), rowptr.data(), rp.data(), rp.size() * sizeof(int32_t)), "upload rowptr");
      throw_on(trl_hip_memcpy_h2d(ctx.get(), col.data(), c.data(), c.size() * sizeof(int32_t)), "upload col");
      std::vector<PanelScalar> vv(v.begin(), v.end());
      throw_on(trl_hip_memcpy_h2d(ctx.get(), val.data(), vv.data(), vv.size() * sizeof(PanelScalar)), "upload val");
      if (m) {
        if (m->size() != n) throw std::invalid_argument("mass size");
        mass = DeviceBuffer<PanelScalar>(ctx, n);
        minv = DeviceBuffer<PanelScalar>(ctx, n);
        std::vector<PanelScalar> mm(n), mi(n);
        for (std::size_t i = 0; i < n; ++i) {
          mm[i] = PanelScalar((*m)[i]);
          mi[i] = PanelScalar(1.0 / (*m)[i]);
        }
        throw_on(trl_hip_memcpy_h2d(ctx.get(), mass.data(), mm.data(), n * sizeof(PanelScalar)), "upload mass");
        throw_on(trl_hip_memcpy_h2d(ctx.get(), minv.data(), mi.data(), n * sizeof(PanelScalar)), "upload mass");
      }
    }
  };

  const int32_t* rowptr_;
  const int32_t* col_;
  const PanelScalar* val_;
  const PanelScalar* minv_;
  std::shared_ptr<SparseOperator<PanelScalar>> op_;
  std::shared_ptr<Owned> owned_;
};

// The generalized problem A x = lambda B x with a sparse symmetric positive
// definite B, by the shift-invert transformation of CSRGeneralizedEVP
// (examples/openmp/csrevp.hh:51-104): apply is Y = (A - sigma B)^{-1} B X,
// dot the B-inner product V^T B W, orthonormalize CholQR in that inner
// product; the Ritz values are mu = 1 / (lambda - sigma), and lambda =
// sigma + 1/mu (the mapping of examples/openmp/gevp_test.cc:135-141).
//
// The reference factorises A - sigma B once (Eigen SparseLU) and solves with
// the factors in every apply.  Here the solve is the device's
// Jacobi-preconditioned conjugate gradient (trl_hip_pcg) over K = A - sigma B
// assembled once, so K must be positive definite: sigma below the smallest
// eigenvalue (for instance sigma = 0 for an SPD A), the setting of the
// reference's example (a stiffness/mass pair, shift 1e-3).  K is checked to
// have a positive diagonal; a direction with p^T K p <= 0 during a solve
// raises std::runtime_error.  The inner tolerance (relative residual of each
// column, default 1e-12) bounds how far the transformed operator is from
// exactly self-adjoint, so it should sit well below the outer tolerance.
// FP64 panels only.
template <unsigned int block_size>
class CSRGeneralizedEVP : public EVPBase<double, block_size> {
public:
  using Base = EVPBase<double, block_size>;
  using BlockView = typename Base::BlockView;
  using BlockMatrixBlockView = typename Base::BlockMatrixBlockView;
  using BlockMultivector = typename Base::BlockMultivector;

  // Host CSR arrays of A and B (full symmetric storage, columns ascending in
  // each row), both n x n, and the shift sigma.
  CSRGeneralizedEVP(const Context& ctx, std::size_t n, const std::vector<int32_t>& a_rowptr, const std::vector<int32_t>& a_col,
                    const std::vector<double>& a_val, const std::vector<int32_t>& b_rowptr, const std::vector<int32_t>& b_col,
                    const std::vector<double>& b_val, double shift)
      : Base(ctx, n), shift_(shift)
  {
    check_csr(n, a_rowptr, a_col, a_val, "A");
    check_csr(n, b_rowptr, b_col, b_val, "B");
    // K = A - sigma B on the union of the two patterns (rows merged in column order)
    std::vector<int32_t> krp(n + 1, 0), kc;
    std::vector<double> kv, dinv(n, 0.0);
    kc.reserve(a_col.size() + b_col.size());
    kv.reserve(a_col.size() + b_col.size());
    for (std::size_t r = 0; r < n; ++r) {
      int32_t i = a_rowptr[r], j = b_rowptr[r];
      const int32_t ie = a_rowptr[r + 1], je = b_rowptr[r + 1];
      while (i < ie || j < je) {
        const int32_t ca = i < ie ? a_col[i] : INT32_MAX, cb = j < je ? b_col[j] : INT32_MAX;
        const int32_t c = std::min(ca, cb);
        double v = 0.0;
        if (ca == c) v += a_val[i++];
        if (cb == c) v -= shift * b_val[j++];
        kc.push_back(c);
        kv.push_back(v);
        if (std::size_t(c) == r) dinv[r] = v;
      }
      krp[r + 1] = int32_t(kc.size());
    }
    for (std::size_t r = 0; r < n; ++r) {
      if (!(dinv[r] > 0.0)) throw std::invalid_argument("CSRGeneralizedEVP: A - shift B has a diagonal entry <= 0 (not positive definite)");
      dinv[r] = 1.0 / dinv[r];
    }
    b_ = std::make_shared<Csr>(ctx, b_rowptr, b_col, b_val);
    k_ = std::make_shared<Csr>(ctx, krp, kc, kv);
    dinv_ = DeviceBuffer<double>(ctx, std::max<std::size_t>(n, 1));
    throw_on(trl_hip_memcpy_h2d(ctx.get(), dinv_.data(), dinv.data(), n * sizeof(double)), "upload 1/diag(K)");
    bop_ = std::make_shared<SparseOperator<double>>(ctx, n, b_->rowptr.data(), b_->col.data(), b_->val.data());
    kop_ = std::make_shared<SparseOperator<double>>(ctx, n, k_->rowptr.data(), k_->col.data(), k_->val.data());
    tmp_ = DeviceBuffer<double>(ctx, std::max<std::size_t>(n * block_size, 1));
    work_ = DeviceBuffer<char>(ctx, std::max<std::size_t>(trl_hip_pcg_workspace(n, int(block_size)), 1));
  }

  // Y = (A - sigma B)^{-1} B X (csrevp.hh:76-84)
  void apply(BlockView X, BlockView Y) override
  {
    bop_->apply(int(block_size), X.data_, tmp_.data());
    int it = 0;
    double rel = 0.0;
    throw_on(trl_hip_pcg(this->ctx_.get(), TRL_HIP_F64, this->N, int(block_size), &kop_->descriptor(), dinv_.data(), tmp_.data(), Y.data_, inner_tol_,
                         inner_maxit_, work_.data(), &it, &rel),
             "shift-invert solve");
    ++solves_;
    inner_iterations_ += unsigned(it);
    max_inner_iterations_ = std::max(max_inner_iterations_, unsigned(it));
  }

  // R = V^T B W (csrevp.hh:86-95)
  void dot(BlockView V, BlockView W, BlockMatrixBlockView R) override
  {
    bop_->apply(int(block_size), W.data_, tmp_.data());
    throw_on(trl_hip_gram(this->ctx_.get(), TRL_HIP_F64, this->N, block_size, 1, V.data_, tmp_.data(), nullptr, R.data_), "B-dot");
  }

  // CholQR in the B-inner product (evp_base.hh:39-71 with the dot above)
  void orthonormalize(BlockView V, BlockMatrixBlockView R) override
  {
    bop_->apply(int(block_size), V.data_, tmp_.data());
    throw_on(trl_hip_cholqr_bprod(this->ctx_.get(), TRL_HIP_F64, this->N, block_size, V.data_, tmp_.data(), R.data_, this->cholqr_passes_),
             "orthonormalize");
  }

  // The multi-block products of the block Gram-Schmidt hooks in the B-inner
  // product: C = V[0:count]^T (B W).
  double* gram_blocks(const BlockMultivector& V, unsigned count, const BlockView& W) override
  {
    double* C = this->coef(std::size_t(count) * block_size * block_size);
    bop_->apply(int(block_size), W.data_, tmp_.data());
    throw_on(trl_hip_gram(this->ctx_.get(), TRL_HIP_F64, this->N, block_size, int(count), V.data(), tmp_.data(), nullptr, C), "gram_blocks");
    return C;
  }
  // Two passes always (the device-side DGKS test needs a diagonal B).
  void reorth_blocks(const BlockMultivector& V, unsigned count, const BlockView& W, double /*eta*/) override
  {
    for (int pass = 0; pass < 2; ++pass) this->update_blocks(V, count, gram_blocks(V, count, W), W);
  }

  double shift() const { return shift_; }
  // lambda = sigma + 1/mu of the current Ritz values mu
  std::vector<double> back_transformed() const
  {
    std::vector<double> l(this->eigenvalues.size());
    for (std::size_t i = 0; i < l.size(); ++i) l[i] = shift_ + 1.0 / this->eigenvalues[i];
    return l;
  }

  void set_inner_tolerance(double t) { inner_tol_ = t; }
  void set_inner_max_iterations(int m) { inner_maxit_ = m; }
  unsigned inner_solves() const { return solves_; }
  unsigned inner_iterations() const { return inner_iterations_; }
  unsigned max_inner_iterations() const { return max_inner_iterations_; }
  const SparseOperator<double>& b_operator() const { return *bop_; }

private:
  struct Csr {
    DeviceBuffer<int32_t> rowptr, col;
    DeviceBuffer<double> val;
    Csr(const Context& ctx, const std::vector<int32_t>& rp, const std::vector<int32_t>& c, const std::vector<double>& v)
        : rowptr(ctx, rp.size()), col(ctx, std::max<std::size_t>(c.size(), 1)), val(ctx, std::max<std::size_t>(v.size(), 1))
    {
      throw_on(trl_hip_memcpy_h2d(ctx.get(), rowptr.data(), rp.data(), rp.size() * sizeof(int32_t)), "upload rowptr");
      throw_on(trl_hip_memcpy_h2d(ctx.get(), col.data(), c.data(), c.size() * sizeof(int32_t)), "upload col");
      throw_on(trl_hip_memcpy_h2d(ctx.get(), val.data(), v.data(), v.size() * sizeof(double)), "upload val");
    }
  };

  static void check_csr(std::size_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& c, const std::vector<double>& v, const char* name)
  {
    if (rp.size() != n + 1 || rp[0] != 0 || c.size() != v.size() || std::size_t(rp[n]) != c.size())
      throw std::invalid_argument(std::string("CSRGeneralizedEVP: inconsistent CSR arrays of ") + name);
    for (std::size_t r = 0; r < n; ++r) {
      if (rp[r + 1] < rp[r]) throw std::invalid_argument(std::string("CSRGeneralizedEVP: row pointers of ") + name + " decrease");
      for (int32_t p = rp[r]; p < rp[r + 1]; ++p)
        if (c[p] < 0 || std::size_t(c[p]) >= n || (p > rp[r] && c[p] <= c[p - 1]))
          throw std::invalid_argument(std::string("CSRGeneralizedEVP: columns of ") + name + " out of range or not ascending");
    }
  }

  double shift_;
  std::shared_ptr<Csr> b_, k_;
  DeviceBuffer<double> dinv_, tmp_;
  DeviceBuffer<char> work_;
  std::shared_ptr<SparseOperator<double>> bop_, kop_;
  double inner_tol_ = 1e-12;
  int inner_maxit_ = 10000;
  unsigned solves_ = 0, inner_iterations_ = 0, max_inner_iterations_ = 0;
};

// Block classical Gram-Schmidt, two passes, through the reference's
// Reorth hook (lanczos.hh:311, reorthogonalization.hh:29-32): each pass is
// one multi-block Gram C = V[0:count]^T M V_next and one multi-block update
// V_next -= V[0:count] C, in place of ModifiedGS's 2*count single-block
// calls (reorthogonalization.hh:14-21).
struct BlockCGS2 {
  template <typename EVP, typename BMV>
  void operator()(EVP& evp, BMV& V, unsigned int count, typename BMV::BlockView V_next, typename BMV::BlockMatrix::BlockView /*tmp*/) const
  {
    for (int pass = 0; pass < 2; ++pass) {
      const double* C = evp.gram_blocks(V, count, V_next);
      evp.update_blocks(V, count, C, V_next);
    }
  }
};

// Block classical Gram-Schmidt with a second pass only where the first one
// cancelled (the DGKS criterion: a column whose norm fell below eta times its
// norm before the pass is orthogonalised again), through the same Reorth hook.
// In the Lanczos recurrence the new block is nearly orthogonal to the basis
// already, so the second pass, and half of the reorthogonalisation traffic,
// is usually skipped.
struct BlockCGSAdaptive {
  double eta = 0.70710678118654752; // 1/sqrt(2)
  template <typename EVP, typename BMV>
  void operator()(EVP& evp, BMV& V, unsigned int count, typename BMV::BlockView V_next, typename BMV::BlockMatrix::BlockView /*tmp*/) const
  {
    evp.reorth_blocks(V, count, V_next, eta);
  }
};

} // namespace trl::hip
