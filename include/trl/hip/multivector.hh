#pragma once
// Block multivectors and block views on device memory.
//
// Mirrors impl/openmp/multivector.hh:20-98 and blockview.hh:20-207 of the
// reference: a multivector owns cols/b contiguous panels, each an n x b
// row-major block (element (r, c) of block i at data + i*n*b + r*b + c),
// zero-initialised, movable and not copyable; a BlockView is a span-like
// non-owning view of one panel.  Panels are FP32 or FP64 (ScalarT); the
// b x b coefficient blocks they are multiplied with are always FP64.

#include <cstddef>
#include <stdexcept>
#include <string>

#include "blockmatrix.hh"
#include "context.hh"

namespace trl::hip {

template <class ScalarT, unsigned int block_size>
class BlockView {
public:
  static constexpr auto blocksize = block_size;
  using EntryType = ScalarT;
  using MatrixBlockView = BlockMatrixBlockView<block_size>;
  static constexpr trl_hip_dtype dtype = dtype_of<ScalarT>();

  BlockView(ScalarT* data, std::size_t n, const Context& ctx) : data_(data), n_(n), ctx_(ctx) {}

  // X = 0 (blockview.hh:31-36)
  void set_zero() { throw_on(trl_hip_fill_zero(ctx_.get(), data_, n_ * block_size * sizeof(ScalarT)), "set_zero"); }

  std::size_t rows() const { return n_; }
  constexpr std::size_t cols() const { return block_size; }

  // X = Y (blockview.hh:59-63)
  void copy_from(const BlockView& other) { throw_on(trl_hip_memcpy_d2d(ctx_.get(), data_, other.data_, n_ * block_size * sizeof(ScalarT)), "copy_from"); }

  // X -= Y (blockview.hh:69-74)
  BlockView& operator-=(const BlockView& other)
  {
    throw_on(trl_hip_panel_sub(ctx_.get(), dtype, n_, block_size, data_, other.data_), "operator-=");
    return *this;
  }

  // other += this * W (blockview.hh:81-97)
  void mult_add(const MatrixBlockView& W, const BlockView& other)
  {
    throw_on(trl_hip_panel_mult(ctx_.get(), dtype, n_, block_size, data_, W.data_, other.data_, TRL_HIP_MULT_ADD), "mult_add");
  }

  // other = this * W (blockview.hh:104-119)
  void mult(const MatrixBlockView& W, const BlockView& other)
  {
    throw_on(trl_hip_panel_mult(ctx_.get(), dtype, n_, block_size, data_, W.data_, other.data_, TRL_HIP_MULT), "mult");
  }

  // other = this * W^T (blockview.hh:126-141)
  void mult_transpose(const MatrixBlockView& W, const BlockView& other)
  {
    throw_on(trl_hip_panel_mult(ctx_.get(), dtype, n_, block_size, data_, W.data_, other.data_, TRL_HIP_MULT_TRANSPOSE), "mult_transpose");
  }

  // R = this^T other (blockview.hh:152-182)
  void dot(const BlockView& other, const MatrixBlockView& R)
  {
    throw_on(trl_hip_gram(ctx_.get(), dtype, n_, block_size, 1, data_, other.data_, nullptr, R.data_), "dot");
  }

  // this -= other * R (blockview.hh:189-201)
  void subtract_product(const BlockView& other, const MatrixBlockView& R)
  {
    throw_on(trl_hip_panel_mult(ctx_.get(), dtype, n_, block_size, data_, R.data_, other.data_, TRL_HIP_SUBTRACT_PRODUCT), "subtract_product");
  }

  // this -= other * R^T in one pass (no counterpart in the concepts: the
  // driver's mult_transpose into a scratch block followed by -=)
  void subtract_product_transpose(const BlockView& other, const MatrixBlockView& R)
  {
    throw_on(trl_hip_panel_mult(ctx_.get(), dtype, n_, block_size, data_, R.data_, other.data_, TRL_HIP_SUBTRACT_PRODUCT_TRANSPOSE),
             "subtract_product_transpose");
  }

  const Context& context() const { return ctx_; }

  ScalarT* data_;

private:
  std::size_t n_;
  Context ctx_;
};

template <class ScalarT, unsigned int block_size>
class BlockMultivector {
public:
  using Scalar = ScalarT;
  using BlockView = ::trl::hip::BlockView<ScalarT, block_size>;
  using BlockMatrix = ::trl::hip::BlockMatrix<block_size>;
  using MatrixBlockView = typename BlockMatrix::BlockView;

  static constexpr auto blocksize = block_size;

  // multivector.hh:38-49: cols must be a multiple of the block size
  BlockMultivector(const Context& ctx, std::size_t rows, std::size_t cols) : ctx_(ctx), rows_(rows), blocks_(cols / block_size)
  {
    if (cols % block_size != 0) throw std::invalid_argument("Number of columns must be divisible by blocksize");
    buf_ = DeviceBuffer<ScalarT>(ctx, rows * cols);
  }

  BlockMultivector(const BlockMultivector&) = delete;
  BlockMultivector& operator=(const BlockMultivector&) = delete;
  BlockMultivector(BlockMultivector&&) = default;
  BlockMultivector& operator=(BlockMultivector&&) = default;

  // multivector.hh:85-89
  BlockView block_view(std::size_t i) const
  {
    if (i >= blocks_) throw std::out_of_range("Block " + std::to_string(i) + " is out of bounds (number of blocks is " + std::to_string(blocks_) + ")");
    return BlockView(buf_.data() + i * rows_ * block_size, rows_, ctx_);
  }

  std::size_t blocks() const { return blocks_; }
  std::size_t rows() const { return rows_; }
  ScalarT* data() const { return buf_.data(); }
  const Context& context() const { return ctx_; }

private:
  Context ctx_;
  std::size_t rows_;
  std::size_t blocks_;
  DeviceBuffer<ScalarT> buf_;
};

} // namespace trl::hip
