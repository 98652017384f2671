#pragma once
// Sparse eigenproblems on the HIP backend.
//
// CSREVP follows examples/openmp/csrevp.hh:19-48 (Y = A X for a symmetric
// CSR matrix); with a diagonal mass M it is the generalized problem
// K x = lambda M x in the form A = M^{-1} K, self-adjoint in the M-inner
// product (the B-inner product of examples/openmp/csrevp.hh:89-97 with a
// diagonal B; no factorisation is needed).  Arrays are device arrays of the
// panel type, either borrowed from the caller or uploaded and owned.

#include <cstdint>
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
