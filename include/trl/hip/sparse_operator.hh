#pragma once
// A device-resident sparse operator Y = diag(d) A X with its kernel choice.
//
// The reference's CSR operator is a fixed loop over rows
// (examples/openmp/csrevp.hh:36-44; one work-item per row in the SYCL form,
// examples/sycl/csrevp.hh:159-187).  Here the same product has several
// kernels (one row per lane, column-split rows, nonzero lane groups) and two
// storage forms (CSR, slot-major ELL), visited in index order or plane block
// by plane block.  tune() times them for one block size on this device and
// keeps the fastest, for the plain product and for the fused axpby step of
// a Chebyshev filter; tune_panels() does the same for that step on several
// panels per launch.  The ELL copy tune() may build is owned by the
// operator.  The CSR arrays are borrowed.

#include <cstdint>
#include <memory>

#include "context.hh"

namespace trl::hip {

template <class S>
class SparseOperator {
public:
  static constexpr trl_hip_dtype dtype = dtype_of<S>();

  // Borrows rowptr[n+1], col[nnz], val[nnz] (device) and the optional row
  // scaling dscale[n]; variant as in trl_hip_spmat (0 = untuned default).
  SparseOperator(const Context& ctx, std::size_t n, const int32_t* rowptr, const int32_t* col, const S* val, const S* dscale = nullptr, int variant = 0)
      : ctx_(ctx), n_(n), csr_{0, rowptr, 0, col, val, dscale, variant, 0, 0, 0, 0, 0}, op_(csr_), op_axpby_(csr_), op_panels_(csr_)
  {
  }

  // Times every kernel for block size b, on CSR and, when the longest row
  // pads the ELL form by at most max_padding (slots / nnz - 1), on ELL, and
  // keeps the fastest, for the plain product and (axpby) the fused axpby
  // form.  Returns the plain product's time per call in microseconds.
  double tune(int b, bool allow_ell = true, double max_padding = 0.25, bool axpby = true)
  {
    if (n_ == 0) return 0.0;
    const trl_hip_spmat* ell = nullptr;
    if (allow_ell) {
      int width = 0;
      throw_on(trl_hip_csr_max_row(ctx_.get(), n_, csr_.rowptr, &width), "csr_max_row");
      int32_t nnz = 0;
      throw_on(trl_hip_memcpy_d2h(ctx_.get(), &nnz, csr_.rowptr + n_, sizeof(int32_t)), "nnz");
      if (width > 0 && double(width) * double(n_) <= (1.0 + max_padding) * double(nnz)) {
        make_ell(width);
        ell = &ell_desc_;
      }
    }
    double us = 0.0;
    throw_on(trl_hip_spmm_tune(ctx_.get(), dtype, n_, b, &csr_, ell, &op_, &us), "spmm_tune");
    if (axpby)
      throw_on(trl_hip_spmm_tune_axpby(ctx_.get(), dtype, n_, b, &csr_, ell, &op_axpby_, &axpby_us_), "spmm_tune_axpby");
    else {
      op_axpby_ = op_; // untimed: the plain product's choice
      axpby_us_ = 0.0;
    }
    if (op_.format == 0 && op_axpby_.format == 0) ell_.reset(); // the ELL copy is only kept while it is used
    tuned_b_ = b;
    tuned_us_ = us;
    panels_n_ = 0;
    return us;
  }

  // Takes another operator's choice of storage and kernel (format, ELL
  // width, variant, row order; e.g. the FP32 copy's, tuned) without timing
  // anything: for an operator that runs too few products to repay a tune.
  // The axpby form takes the same descriptor, untimed (axpby_us() is 0, so
  // panels_pay() stays false until a tune).
  void adopt(int b, const trl_hip_spmat& like)
  {
    if (n_ == 0) return;
    if (like.format == 1) {
      make_ell(like.width);
      op_ = ell_desc_;
    }
    else
      op_ = csr_;
    op_.variant = like.variant;
    op_.plane = like.plane;
    op_.plane_block = like.plane_block;
    op_axpby_ = op_;
    tuned_b_ = b;
    tuned_us_ = axpby_us_ = 0.0;
    panels_n_ = 0;
  }

  // Takes the blocked row order (plane, plane_block) of another operator's
  // descriptor for the one-panel forms, untimed: an FP64 operator that adopted
  // an FP32 operator's choice, once the FP32 panel tune has found blocks of a
  // plane (FP64 rows, twice as wide, gain from them where FP32 rows in index
  // order do not: 175 against 207 us for the one-panel FP64 product of the
  // 160^3 stencil).  No change for a descriptor in index order.
  void adopt_row_order(const trl_hip_spmat& like)
  {
    if (n_ == 0 || like.plane <= 0) return;
    op_.plane = op_axpby_.plane = like.plane;
    op_.plane_block = op_axpby_.plane_block = like.plane_block;
  }

  // After tune(b): times the axpby form on npanels panels at once (the
  // matrix read once per launch) for axpby_panels.  Returns the time per
  // launch in microseconds.
  double tune_panels(int b, int npanels)
  {
    if (n_ == 0 || npanels < 2) return 0.0;
    const trl_hip_spmat* ell = ell_ ? &ell_desc_ : nullptr;
    throw_on(trl_hip_spmm_tune_panels(ctx_.get(), dtype, n_, b, npanels, &csr_, ell, &op_panels_, &panels_us_), "spmm_tune_panels");
    panels_n_ = npanels;
    return panels_us_;
  }

  // True when tune_panels(b, npanels) found one launch over the panels
  // faster than npanels one-panel launches.
  bool panels_pay(int npanels) const { return panels_n_ == npanels && panels_us_ < 0.97 * npanels * axpby_us_; }

  // Y = diag(dscale) A X (scaled = false: Y = A X).
  void apply(int b, const S* X, S* Y, bool scaled = true) const
  {
    trl_hip_spmat d = op_;
    if (!scaled) d.dscale = nullptr;
    throw_on(trl_hip_spmm(ctx_.get(), dtype, n_, b, &d, X, Y), "apply");
  }

  // Y = alpha diag(dscale) A X + beta X + gamma Z (with the variant tuned for this form).
  void axpby(int b, const S* X, S* Y, const S* Z, double alpha, double beta, double gamma) const
  {
    throw_on(trl_hip_spmm_axpby(ctx_.get(), dtype, n_, b, &op_axpby_, X, Y, Z, alpha, beta, gamma), "apply_axpby");
  }

  // The same on npanels consecutive panels (panel j at j*n*b).
  void apply_panels(int b, int npanels, const S* X, S* Y) const
  {
    throw_on(trl_hip_spmm_panels(ctx_.get(), dtype, n_, b, &op_, npanels, X, Y), "apply_panels");
  }
  void axpby_panels(int b, int npanels, const S* X, S* Y, const S* Z, double alpha, double beta, double gamma) const
  {
    const trl_hip_spmat& d = panels_n_ > 1 && npanels > 1 ? op_panels_ : op_axpby_; // the panel form, tuned at panels_n_ panels
    throw_on(trl_hip_spmm_axpby_panels(ctx_.get(), dtype, n_, b, &d, npanels, X, Y, Z, alpha, beta, gamma), "apply_axpby_panels");
  }
  // The same plus the forcing term F_j diag(g_j) (g: device, npanels*b).
  void axpby_forced(int b, int npanels, const S* X, S* Y, const S* Z, double alpha, double beta, double gamma, const S* F, const double* g) const
  {
    const trl_hip_spmat& d = panels_n_ > 1 && npanels > 1 ? op_panels_ : op_axpby_;
    throw_on(trl_hip_spmm_axpby_forced(ctx_.get(), dtype, n_, b, &d, npanels, X, Y, Z, alpha, beta, gamma, F, g), "apply_axpby_forced");
  }

  const trl_hip_spmat& descriptor() const { return op_; }
  const trl_hip_spmat& axpby_descriptor() const { return op_axpby_; }
  const trl_hip_spmat& panels_descriptor() const { return op_panels_; }
  double axpby_us() const { return axpby_us_; }
  double panels_us() const { return panels_us_; }
  std::size_t size() const { return n_; }
  int tuned_block_size() const { return tuned_b_; }
  double tuned_us() const { return tuned_us_; }

private:
  struct Ell {
    int width;
    DeviceBuffer<int32_t> col;
    DeviceBuffer<S> val;
    Ell(const Context& c, std::size_t n, int w) : width(w), col(c, std::size_t(w) * n), val(c, std::size_t(w) * n) {}
  };

  // the slot-major ELL copy of the CSR arrays with `width` slots
  void make_ell(int width)
  {
    if (!ell_ || ell_->width != width) {
      ell_ = std::make_shared<Ell>(ctx_, n_, width);
      throw_on(trl_hip_csr_to_ell(ctx_.get(), dtype, n_, csr_.rowptr, csr_.col, csr_.val, width, ell_->col.data(), ell_->val.data()), "csr_to_ell");
    }
    ell_desc_ = trl_hip_spmat{1, nullptr, width, ell_->col.data(), ell_->val.data(), csr_.dscale, 0, 0, 0, 0, 0, 0};
  }

  Context ctx_;
  std::size_t n_;
  trl_hip_spmat csr_;
  trl_hip_spmat op_;
  trl_hip_spmat op_axpby_;
  trl_hip_spmat op_panels_;
  trl_hip_spmat ell_desc_{};
  std::shared_ptr<Ell> ell_;
  int tuned_b_ = 0;
  double tuned_us_ = 0.0, axpby_us_ = 0.0, panels_us_ = 0.0;
  int panels_n_ = 0;
};

} // namespace trl::hip
