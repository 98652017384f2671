// Sparse operator kernels: Y = diag(d) A X with X, Y n x b row-major panels.
//
// CSR: L lanes of a wavefront share one row (L in {1,2,4,8,16}).  With L = 1
// a lane owns a row and loads its indices/values in chunks before gathering
// the b-wide rows of X into registers; with L > 1 the lanes of a group walk
// the row's nonzeros with coalesced index/value loads and the partial rows are
// combined by a butterfly (shfl_xor) reduction inside the group.  ELL: one row
// per lane, slot-major arrays (fully coalesced), chunked the same way.
//
// Workgroups are renumbered so that each XCD walks a contiguous range of rows:
// the stencil neighbours of a row (r +- 1, r +- nx, ...) then share that XCD's L2.
#include "common.hh"

using namespace trl_hip_dev;

namespace {

constexpr int kThreads = 256;
constexpr unsigned kXcds = 8;

__device__ inline unsigned xcd_block(unsigned bid, unsigned nb)
{
  if (nb % kXcds != 0) return bid;
  return (bid % kXcds) * (nb / kXcds) + bid / kXcds;
}

// Nonzeros whose indices and values are loaded before any of their gathers.
constexpr int kChunk = 8;

// Accumulates row `row` of A X into acc.  L == 1: one lane per row, the
// row's column indices and values are loaded kChunk at a time before the
// gathers of X (two memory latencies per chunk instead of two per nonzero);
// padding slots gather the row's own X row with value 0.  L > 1: lanes
// sub, sub+L, ... of the row, combined by a butterfly in the lane group.
template <typename S, int B, int L>
__device__ inline void csr_row(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const S* __restrict__ val, const S* __restrict__ X, int row,
                               int sub, S (&acc)[B])
{
#pragma unroll
  for (int c = 0; c < B; ++c) acc[c] = S(0);
  const int beg = rowptr[row], end = rowptr[row + 1];
  if constexpr (L == 1) {
    for (int p0 = beg; p0 < end; p0 += kChunk) {
      int ci[kChunk];
      S vi[kChunk];
#pragma unroll
      for (int q = 0; q < kChunk; ++q) {
        const bool ok = p0 + q < end;
        ci[q] = ok ? col[p0 + q] : row;
        vi[q] = ok ? val[p0 + q] : S(0);
      }
      S x[kChunk][B];
#pragma unroll
      for (int q = 0; q < kChunk; ++q) RowIO<S, B>::load(X + size_t(ci[q]) * B, x[q]);
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
#pragma unroll
        for (int c = 0; c < B; ++c) acc[c] = fma(vi[q], x[q][c], acc[c]);
    }
  }
  else {
    for (int p = beg + sub; p < end; p += L) {
      const int c = col[p];
      const S v = val[p];
      S x[B];
      RowIO<S, B>::load(X + size_t(c) * B, x);
#pragma unroll
      for (int q = 0; q < B; ++q) acc[q] = fma(v, x[q], acc[q]);
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1)
#pragma unroll
      for (int q = 0; q < B; ++q) acc[q] += __shfl_xor(acc[q], off, L);
  }
}

template <typename S, int B, int L>
__global__ __launch_bounds__(kThreads) void k_spmm_csr(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const S* __restrict__ val,
                                                       const S* __restrict__ X, S* __restrict__ Y, const S* __restrict__ dscale, int n)
{
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int gid = int(blk * blockDim.x + threadIdx.x);
  const int row = gid / L;
  const int sub = int(threadIdx.x) % L;
  if (row >= n) return; // whole lane groups leave together
  S acc[B];
  csr_row<S, B, L>(rowptr, col, val, X, row, sub, acc);
  if (sub != 0) return;
  if (dscale) {
    const S d = dscale[row];
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] *= d;
  }
  RowIO<S, B>::store(Y + size_t(row) * B, acc);
}

// Fused three-term form: Y = alpha D A X + beta X + gamma Z (D = diag(dscale)
// or I; Z may be Y, or NULL), one Chebyshev-recurrence step per launch.
template <typename S, int B, int L>
__global__ __launch_bounds__(kThreads) void k_spmm_csr_axpby(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const S* __restrict__ val,
                                                             const S* __restrict__ X, S* Y, const S* Z, const S* __restrict__ dscale, double alpha,
                                                             double beta, double gamma, int n)
{
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int gid = int(blk * blockDim.x + threadIdx.x);
  const int row = gid / L;
  const int sub = int(threadIdx.x) % L;
  if (row >= n) return;
  S acc[B];
  csr_row<S, B, L>(rowptr, col, val, X, row, sub, acc);
  if (sub != 0) return;
  S xr[B], zr[B], out[B];
  RowIO<S, B>::load(X + size_t(row) * B, xr);
  if (Z) RowIO<S, B>::load(Z + size_t(row) * B, zr);
  const double a = dscale ? alpha * double(dscale[row]) : alpha;
#pragma unroll
  for (int q = 0; q < B; ++q) {
    double v = a * double(acc[q]) + beta * double(xr[q]);
    if (Z) v += gamma * double(zr[q]);
    out[q] = S(v);
  }
  RowIO<S, B>::store(Y + size_t(row) * B, out);
}

template <typename S, int B>
__global__ __launch_bounds__(kThreads) void k_spmm_ell(int width, const int32_t* __restrict__ col, const S* __restrict__ val, const S* __restrict__ X,
                                                       S* __restrict__ Y, const S* __restrict__ dscale, int n)
{
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int row = int(blk * blockDim.x + threadIdx.x);
  if (row >= n) return;
  S acc[B];
#pragma unroll
  for (int c = 0; c < B; ++c) acc[c] = S(0);
  for (int s0 = 0; s0 < width; s0 += kChunk) {
    int ci[kChunk];
    S vi[kChunk];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      const bool ok = s0 + q < width;
      const size_t p = size_t(s0 + q) * n + row;
      ci[q] = ok ? col[p] : row;
      vi[q] = ok ? val[p] : S(0);
    }
    S x[kChunk][B];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) RowIO<S, B>::load(X + size_t(ci[q]) * B, x[q]);
#pragma unroll
    for (int q = 0; q < kChunk; ++q)
#pragma unroll
      for (int c = 0; c < B; ++c) acc[c] = fma(vi[q], x[q][c], acc[c]);
  }
  if (dscale) {
    const S d = dscale[row];
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] *= d;
  }
  RowIO<S, B>::store(Y + size_t(row) * B, acc);
}

template <typename S, int B, int L>
int launch_csr_l(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, const S* val, const S* X, S* Y, const S* ds)
{
  const size_t threads = n * L;
  size_t g = (threads + kThreads - 1) / kThreads;
  // round the grid up to a multiple of the XCD count so the renumbering applies
  g = (g + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL((k_spmm_csr<S, B, L>), dim3(unsigned(g)), dim3(kThreads), 0, ctx->stream, rowptr, col, val, X, Y, ds, int(n));
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

template <typename S, int B, int L>
int launch_axpby_l(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, const S* val, const S* X, S* Y, const S* Z, const S* ds, double a,
                   double b, double g)
{
  size_t grid = (n * L + kThreads - 1) / kThreads;
  grid = (grid + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL((k_spmm_csr_axpby<S, B, L>), dim3(unsigned(grid)), dim3(kThreads), 0, ctx->stream, rowptr, col, val, X, Y, Z, ds, a, b, g, int(n));
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

} // namespace

namespace trl_hip_impl {

int launch_spmm_axpby(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const int32_t* rowptr, const int32_t* col, const void* val, const void* dscale,
                      const void* X, void* Y, const void* Z, double alpha, double beta, double gamma)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const S* v = static_cast<const S*>(val);
      const S* x = static_cast<const S*>(X);
      S* y = static_cast<S*>(Y);
      const S* z = static_cast<const S*>(Z);
      const S* ds = static_cast<const S*>(dscale);
      if constexpr (B >= 8) return launch_axpby_l<S, B, 1>(ctx, n, rowptr, col, v, x, y, z, ds, alpha, beta, gamma);
      else return launch_axpby_l<S, B, 4>(ctx, n, rowptr, col, v, x, y, z, ds, alpha, beta, gamma);
    });
  });
  return 0;
}

int launch_spmm_csr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const int32_t* rowptr, const int32_t* col, const void* val, const void* X, void* Y,
                    const void* dscale, int lanes)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const S* v = static_cast<const S*>(val);
      const S* x = static_cast<const S*>(X);
      S* y = static_cast<S*>(Y);
      const S* d = static_cast<const S*>(dscale);
      int L = lanes;
      if (L == 0) L = (B >= 8) ? 1 : 4;
      switch (L) {
      case 1:
        return launch_csr_l<S, B, 1>(ctx, n, rowptr, col, v, x, y, d);
      case 2:
        return launch_csr_l<S, B, 2>(ctx, n, rowptr, col, v, x, y, d);
      case 4:
        return launch_csr_l<S, B, 4>(ctx, n, rowptr, col, v, x, y, d);
      case 8:
        return launch_csr_l<S, B, 8>(ctx, n, rowptr, col, v, x, y, d);
      default:
        return launch_csr_l<S, B, 16>(ctx, n, rowptr, col, v, x, y, d);
      }
    });
  });
  return 0;
}

int launch_spmm_ell(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int width, const int32_t* col, const void* val, const void* X, void* Y, const void* dscale)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      size_t g = (n + kThreads - 1) / kThreads;
      g = (g + kXcds - 1) / kXcds * kXcds;
      hipLaunchKernelGGL((k_spmm_ell<S, B>), dim3(unsigned(g)), dim3(kThreads), 0, ctx->stream, width, col, static_cast<const S*>(val),
                         static_cast<const S*>(X), static_cast<S*>(Y), static_cast<const S*>(dscale), int(n));
      TRL_HIP_LAUNCH_CHECK();
    });
  });
  return 0;
}

} // namespace trl_hip_impl
