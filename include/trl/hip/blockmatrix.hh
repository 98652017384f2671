#pragma once
// Small dense blocks of the projected problem (T, beta, Ritz coefficients).
//
// Mirrors impl/openmp/matrixblockview.hh:16-77 and blockmatrix.hh:19-65 of the
// reference: a BlockMatrix is a grid of b x b row-major blocks, block (i, j)
// at (i*block_cols + j)*b*b; a view is a non-owning pointer to one block.
// Here the storage is FP64 device memory for both panel precisions, so the
// projected matrix never leaves the GPU during a Lanczos step; every view
// operation is a queued kernel or copy on the context's stream.

#include <cstddef>
#include <span>
#include <vector>

#include "context.hh"

namespace trl::hip {

template <unsigned int block_size>
class BlockMatrixBlockView {
public:
  using EntryType = double;

  BlockMatrixBlockView(double* data, const Context& ctx) : data_(data), ctx_(ctx) {}

  // A = B (matrixblockview.hh:30-34)
  void copy_from(const BlockMatrixBlockView& other) { throw_on(trl_hip_small_copy(ctx_.get(), block_size, other.data_, data_, 0), "copy_from"); }

  // A = B^T (matrixblockview.hh:36-43)
  void copy_from_transpose(const BlockMatrixBlockView& other)
  {
    throw_on(trl_hip_small_copy(ctx_.get(), block_size, other.data_, data_, 1), "copy_from_transpose");
  }

  // A = 0 (matrixblockview.hh:49)
  void set_zero() { throw_on(trl_hip_small_zero(ctx_.get(), block_size, data_), "set_zero"); }

  // C = A B (matrixblockview.hh:55-64)
  void mult(const BlockMatrixBlockView& B, const BlockMatrixBlockView& C) const
  {
    throw_on(trl_hip_small_mult(ctx_.get(), block_size, data_, B.data_, C.data_), "mult");
  }

  // diag(A) = values, off-diagonal entries kept (matrixblockview.hh:71-74);
  // values live on the host.
  void set_diagonal(std::span<double> values) { throw_on(trl_hip_small_set_diagonal(ctx_.get(), block_size, data_, values.data()), "set_diagonal"); }

  // Host copy of the block (synchronous).
  std::vector<double> to_host() const
  {
    std::vector<double> h(block_size * block_size);
    throw_on(trl_hip_memcpy_d2h(ctx_.get(), h.data(), data_, h.size() * sizeof(double)), "to_host");
    return h;
  }

  double* data_;
  Context ctx_;
};

template <unsigned int block_size>
class BlockMatrix {
public:
  using BlockView = BlockMatrixBlockView<block_size>;

  BlockMatrix(const Context& ctx, std::size_t brows, std::size_t bcols)
      : ctx_(ctx), block_rows_(brows), block_cols_(bcols), buf_(ctx, brows * bcols * block_size * block_size)
  {
  }

  BlockMatrix(const BlockMatrix&) = delete;
  BlockMatrix& operator=(const BlockMatrix&) = delete;
  BlockMatrix(BlockMatrix&&) = default;
  BlockMatrix& operator=(BlockMatrix&&) = default;

  std::size_t block_rows() const { return block_rows_; }
  std::size_t block_cols() const { return block_cols_; }

  BlockView block_view(std::size_t block_row, std::size_t block_col) const
  {
    if (block_row >= block_rows_ || block_col >= block_cols_) throw std::out_of_range("BlockMatrix::block_view index out of range");
    return BlockView(buf_.data() + (block_row * block_cols_ + block_col) * block_size * block_size, ctx_);
  }

  double* data() const { return buf_.data(); }
  std::size_t size() const { return buf_.size(); }
  const Context& context() const { return ctx_; }

  std::vector<double> to_host() const
  {
    std::vector<double> h(buf_.size());
    throw_on(trl_hip_memcpy_d2h(ctx_.get(), h.data(), buf_.data(), h.size() * sizeof(double)), "BlockMatrix::to_host");
    return h;
  }
  void from_host(const std::vector<double>& h)
  {
    if (h.size() != buf_.size()) throw std::invalid_argument("BlockMatrix::from_host size mismatch");
    throw_on(trl_hip_memcpy_h2d(ctx_.get(), buf_.data(), h.data(), h.size() * sizeof(double)), "BlockMatrix::from_host");
  }

private:
  Context ctx_;
  std::size_t block_rows_;
  std::size_t block_cols_;
  DeviceBuffer<double> buf_;
};

} // namespace trl::hip
