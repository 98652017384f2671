// Streaming panel kernels: zero fill, dtype conversion, X -= Y, the four
// b x b panel products of the BlockVectorView concept, the small b x b block
// operations, and the residual norms of Ritz pairs.
//
// All panel kernels are HBM-streaming: one row of b scalars per thread,
// loaded/stored with the widest aligned vector the row allows; b x b
// coefficients are staged once per workgroup in LDS and applied in FP64.
#include "common.hh"

using namespace trl_hip_dev;

namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(size_t work, int per_block, int max_blocks)
{
  size_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > size_t(max_blocks)) g = max_blocks;
  return unsigned(g);
}

__global__ __launch_bounds__(kThreads) void k_fill_zero16(uint4* __restrict__ p, size_t n16)
{
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

__global__ __launch_bounds__(kThreads) void k_fill_zero1(unsigned char* __restrict__ p, size_t n)
{
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

template <typename A, typename B>
__global__ __launch_bounds__(kThreads) void k_convert(const A* __restrict__ src, B* __restrict__ dst, size_t n)
{
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = static_cast<B>(src[i]);
}

template <typename S>
__global__ __launch_bounds__(kThreads) void k_sub(S* __restrict__ X, const S* __restrict__ Y, size_t count)
{
  using V = __attribute__((ext_vector_type(16 / sizeof(S)))) S;
  constexpr int per = 16 / sizeof(S);
  const size_t nv = count / per;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  V* xv = reinterpret_cast<V*>(X);
  const V* yv = reinterpret_cast<const V*>(Y);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nv; i += stride) xv[i] = xv[i] - yv[i];
  // tail (count not a multiple of the vector width)
  const size_t t0 = nv * per;
  const size_t gi = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t0 + gi < count) X[t0 + gi] -= Y[t0 + gi];
}

template <typename S>
__global__ __launch_bounds__(kThreads) void k_sub_scalar(S* __restrict__ X, const S* __restrict__ Y, size_t count)
{
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) X[i] -= Y[i];
}

// mode 0: Y = X W ; 1: Y += X W ; 2: Y = X W^T ; 3: X -= Y W
template <typename S, int B, int MODE>
__global__ __launch_bounds__(kThreads) void k_panel_mult(S* __restrict__ X, const double* __restrict__ Wg, S* __restrict__ Y, size_t n)
{
  __shared__ double Ws[B * B];
  for (int i = threadIdx.x; i < B * B; i += blockDim.x) {
    const int r = i / B, c = i % B;
    Ws[i] = (MODE == 2 || MODE == 4) ? Wg[c * B + r] : Wg[i];
  }
  __syncthreads();
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t row = size_t(blockIdx.x) * blockDim.x + threadIdx.x; row < n; row += stride) {
    S in[B];
    const S* src = (MODE >= 3) ? Y : X;
    RowIO<S, B>::load(src + row * B, in);
    double acc[B];
#pragma unroll
    for (int c = 0; c < B; ++c) acc[c] = 0.0;
#pragma unroll
    for (int cp = 0; cp < B; ++cp) {
      const double x = static_cast<double>(in[cp]);
#pragma unroll
      for (int c = 0; c < B; ++c) acc[c] = fma(x, Ws[cp * B + c], acc[c]);
    }
    S out[B];
    if constexpr (MODE == 0 || MODE == 2) {
#pragma unroll
      for (int c = 0; c < B; ++c) out[c] = static_cast<S>(acc[c]);
      RowIO<S, B>::store(Y + row * B, out);
    }
    else if constexpr (MODE == 1) {
      S old[B];
      RowIO<S, B>::load(Y + row * B, old);
#pragma unroll
      for (int c = 0; c < B; ++c) out[c] = static_cast<S>(static_cast<double>(old[c]) + acc[c]);
      RowIO<S, B>::store(Y + row * B, out);
    }
    else {
      S old[B];
      RowIO<S, B>::load(X + row * B, old);
#pragma unroll
      for (int c = 0; c < B; ++c) out[c] = static_cast<S>(static_cast<double>(old[c]) - acc[c]);
      RowIO<S, B>::store(X + row * B, out);
    }
  }
}

// small b x b ops, one workgroup: op 0 zero, 1 transpose copy, 2 set diagonal, 3 C = A B
struct DiagVals {
  double v[16];
};

__global__ void k_small(int op, int b, const double* __restrict__ A, const double* __restrict__ Bm, double* __restrict__ C, DiagVals d)
{
  const int t = threadIdx.x;
  if (t >= b * b) return;
  const int r = t / b, c = t % b;
  if (op == 0) C[t] = 0.0;
  else if (op == 1) C[t] = A[c * b + r];
  else if (op == 2) {
    if (r == c) C[t] = d.v[r];
  }
  else {
    double s = 0.0;
    for (int k = 0; k < b; ++k) s += A[r * b + k] * Bm[k * b + c];
    C[t] = s;
  }
}

// residual sums: per column c, sum (Y - lambda_c m X)^2 and sum Y^2; two-level
template <typename S, int B>
__global__ __launch_bounds__(kThreads) void k_residual(const S* __restrict__ X, const S* __restrict__ Y, const S* __restrict__ md, DiagVals lam, size_t n,
                                                       double* __restrict__ partials)
{
  double r2[B], y2[B];
#pragma unroll
  for (int c = 0; c < B; ++c) r2[c] = y2[c] = 0.0;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t row = size_t(blockIdx.x) * blockDim.x + threadIdx.x; row < n; row += stride) {
    S x[B], y[B];
    RowIO<S, B>::load(X + row * B, x);
    RowIO<S, B>::load(Y + row * B, y);
    const double m = md ? static_cast<double>(md[row]) : 1.0;
#pragma unroll
    for (int c = 0; c < B; ++c) {
      const double yy = static_cast<double>(y[c]);
      const double rr = yy - lam.v[c] * m * static_cast<double>(x[c]);
      r2[c] = fma(rr, rr, r2[c]);
      y2[c] = fma(yy, yy, y2[c]);
    }
  }
  __shared__ double red[kThreads];
  for (int q = 0; q < 2 * B; ++q) {
    red[threadIdx.x] = q < B ? r2[q] : y2[q - B];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if (threadIdx.x < unsigned(s)) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[size_t(blockIdx.x) * 2 * B + q] = red[0];
    __syncthreads();
  }
}

// Residual panels of Ritz pairs: R_p = AX_p - X_p diag(theta_p) in FP64,
// stored in FP32 as R and, times s, as E (panel p = blockIdx.y).
template <int B>
__global__ __launch_bounds__(kThreads) void k_residual_split(const double* __restrict__ AX, const double* __restrict__ X, const double* __restrict__ theta,
                                                             double s, size_t n, float* __restrict__ R, float* __restrict__ E)
{
  const size_t po = size_t(blockIdx.y) * n * B;
  double th[B];
#pragma unroll
  for (int c = 0; c < B; ++c) th[c] = theta[size_t(blockIdx.y) * B + c];
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += stride) {
    double ax[B], x[B];
    RowIO<double, B>::template load<true>(AX + po + r * B, ax);
    RowIO<double, B>::template load<true>(X + po + r * B, x);
    float rr[B], ee[B];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      const double v = ax[c] - th[c] * x[c];
      rr[c] = float(v);
      ee[c] = float(s * v);
    }
    RowIO<float, B>::template store<true>(R + po + r * B, rr);
    RowIO<float, B>::template store<true>(E + po + r * B, ee);
  }
}

// X_p = X_p diag(h_p) + E_p, E in FP32 (panel p = blockIdx.y).
template <int B>
__global__ __launch_bounds__(kThreads) void k_combine(double* __restrict__ X, const double* __restrict__ h, const float* __restrict__ E, size_t n)
{
  const size_t po = size_t(blockIdx.y) * n * B;
  double hh[B];
#pragma unroll
  for (int c = 0; c < B; ++c) hh[c] = h[size_t(blockIdx.y) * B + c];
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += stride) {
    double x[B];
    float e[B];
    RowIO<double, B>::template load<true>(X + po + r * B, x);
    RowIO<float, B>::template load<true>(E + po + r * B, e);
#pragma unroll
    for (int c = 0; c < B; ++c) x[c] = hh[c] * x[c] + double(e[c]);
    RowIO<double, B>::template store<true>(X + po + r * B, x);
  }
}

// k_residual_split with consecutive lanes on consecutive 16-byte pieces of
// the FP64 panels (two entries of AX and X, the matching two FP32 entries of
// R and E per lane) instead of one row each, so that every load and store
// instruction of a wave covers one contiguous range.  One panel per launch
// (blockIdx.y unused); count = n * B entries.
template <int B>
__global__ __launch_bounds__(kThreads) void k_residual_split_flat(const double* __restrict__ AX, const double* __restrict__ X,
                                                                  const double* __restrict__ theta, double s, size_t count, float* __restrict__ R,
                                                                  float* __restrict__ E)
{
  using D2 = double __attribute__((ext_vector_type(2)));
  using F2 = float __attribute__((ext_vector_type(2)));
  const size_t pairs = count / 2;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < pairs; i += stride) {
    const D2 ax = __builtin_nontemporal_load(reinterpret_cast<const D2*>(AX) + i);
    const D2 x = __builtin_nontemporal_load(reinterpret_cast<const D2*>(X) + i);
    const int c = int((2 * i) % B);
    const double v0 = ax[0] - theta[c] * x[0], v1 = ax[1] - theta[c + 1] * x[1];
    __builtin_nontemporal_store(F2{float(v0), float(v1)}, reinterpret_cast<F2*>(R) + i);
    __builtin_nontemporal_store(F2{float(s * v0), float(s * v1)}, reinterpret_cast<F2*>(E) + i);
  }
}

// The same with consecutive lanes on consecutive 16-byte pieces of the
// panel (two FP64 entries of X and the matching two FP32 entries of E per
// lane) instead of one row each: every load and store instruction of a wave
// covers one contiguous range.  One panel per launch; count = n * B entries.
template <int B>
__global__ __launch_bounds__(kThreads) void k_combine_flat(double* __restrict__ X, const double* __restrict__ h, const float* __restrict__ E, size_t count)
{
  const size_t po = 0;
  const double* hp = h;
  using D2 = double __attribute__((ext_vector_type(2)));
  using F2 = float __attribute__((ext_vector_type(2)));
  const size_t pairs = count / 2;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < pairs; i += stride) {
    D2 x = __builtin_nontemporal_load(reinterpret_cast<const D2*>(X + po) + i);
    const F2 e = __builtin_nontemporal_load(reinterpret_cast<const F2*>(E + po) + i);
    const int c = int((2 * i) % B);
    x[0] = hp[c] * x[0] + double(e[0]);
    x[1] = hp[c + 1] * x[1] + double(e[1]);
    __builtin_nontemporal_store(x, reinterpret_cast<D2*>(X + po) + i);
  }
}

// out[p] = sum over w < nwg of partials[w * P + p]: one wave per column p
// (workgroup p), lanes over the workgroups' partials in a fixed order, then a
// butterfly: every load of the column in flight at once instead of nwg
// dependent steps of one lane.
__global__ __launch_bounds__(64) void k_sum_partials(const double* __restrict__ partials, int nwg, int P, double* __restrict__ out)
{
  const int p = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int w = lane; w < nwg; w += 64) s += partials[size_t(w) * P + p];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[p] = s;
}

} // namespace

namespace trl_hip_impl {

int launch_fill_zero(trl_hip_ctx* ctx, void* dst, size_t bytes)
{
  const uintptr_t a = reinterpret_cast<uintptr_t>(dst);
  if (a % 16 == 0 && bytes % 16 == 0) {
    const size_t n16 = bytes / 16;
    hipLaunchKernelGGL(k_fill_zero16, dim3(grid_for(n16, kThreads, 4096)), dim3(kThreads), 0, ctx->stream, static_cast<uint4*>(dst), n16);
  }
  else {
    hipLaunchKernelGGL(k_fill_zero1, dim3(grid_for(bytes, kThreads, 4096)), dim3(kThreads), 0, ctx->stream, static_cast<unsigned char*>(dst), bytes);
  }
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_residual_split(trl_hip_ctx* ctx, size_t n, int b, int npanels, const double* AX, const double* X, const double* theta, double s, float* R,
                          float* E)
{
  if (b % 2 == 0) {
    // one launch per panel, lanes on consecutive 16-byte pieces (5 FP64 panels
    // of 4M x 8: 721 us against 989 us for the row-per-lane kernel)
    const dim3 g(grid_for(n * size_t(b) / 2, kThreads, 8192)), blk(kThreads);
    for (int j = 0; j < npanels; ++j) {
      const size_t po = size_t(j) * n * size_t(b);
      TRL_DISPATCH_B(b, {
        if constexpr (B % 2 == 0)
          hipLaunchKernelGGL((k_residual_split_flat<B>), g, blk, 0, ctx->stream, AX + po, X + po, theta + size_t(j) * B, s, n * B, R + po, E + po);
      });
    }
  }
  else {
    const dim3 g(grid_for(n, kThreads, 2048), unsigned(npanels)), blk(kThreads);
    TRL_DISPATCH_B(b, { hipLaunchKernelGGL((k_residual_split<B>), g, blk, 0, ctx->stream, AX, X, theta, s, n, R, E); });
  }
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_combine(trl_hip_ctx* ctx, size_t n, int b, int npanels, double* X, const double* h, const float* E)
{
  if (b % 2 == 0) {
    // one launch per panel, lanes on consecutive 16-byte pieces (measured on
    // 5 FP64 panels of 4M x 8: 578 us against 674 us for one launch over all
    // panels and 1184 us for the row-per-lane kernel)
    const dim3 g(grid_for(n * size_t(b) / 2, kThreads, 8192)), blk(kThreads);
    for (int j = 0; j < npanels; ++j) {
      const size_t po = size_t(j) * n * size_t(b);
      TRL_DISPATCH_B(b, {
        if constexpr (B % 2 == 0) hipLaunchKernelGGL((k_combine_flat<B>), g, blk, 0, ctx->stream, X + po, h + size_t(j) * B, E + po, n * B);
      });
    }
  }
  else {
    const dim3 g(grid_for(n, kThreads, 2048), unsigned(npanels)), blk(kThreads);
    TRL_DISPATCH_B(b, { hipLaunchKernelGGL((k_combine<B>), g, blk, 0, ctx->stream, X, h, E, n); });
  }
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_convert(trl_hip_ctx* ctx, trl_hip_dtype st, const void* src, trl_hip_dtype dt, void* dst, size_t count)
{
  const dim3 g(grid_for(count, kThreads, 8192)), blk(kThreads);
  if (st == TRL_HIP_F32 && dt == TRL_HIP_F64)
    hipLaunchKernelGGL((k_convert<float, double>), g, blk, 0, ctx->stream, static_cast<const float*>(src), static_cast<double*>(dst), count);
  else if (st == TRL_HIP_F64 && dt == TRL_HIP_F32)
    hipLaunchKernelGGL((k_convert<double, float>), g, blk, 0, ctx->stream, static_cast<const double*>(src), static_cast<float*>(dst), count);
  else if (st == TRL_HIP_F32 && dt == TRL_HIP_F32)
    hipLaunchKernelGGL((k_convert<float, float>), g, blk, 0, ctx->stream, static_cast<const float*>(src), static_cast<float*>(dst), count);
  else if (st == TRL_HIP_F64 && dt == TRL_HIP_F64)
    hipLaunchKernelGGL((k_convert<double, double>), g, blk, 0, ctx->stream, static_cast<const double*>(src), static_cast<double*>(dst), count);
  else
    return TRL_HIP_EINVAL;
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_panel_sub(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const void* Y)
{
  const size_t count = n * size_t(b);
  TRL_DISPATCH_T(t, {
    constexpr int per = 16 / sizeof(S);
    const uintptr_t ax = reinterpret_cast<uintptr_t>(X), ay = reinterpret_cast<uintptr_t>(Y);
    if (ax % 16 || ay % 16) {
      hipLaunchKernelGGL(k_sub_scalar<S>, dim3(grid_for(count, kThreads, 8192)), dim3(kThreads), 0, ctx->stream, static_cast<S*>(X),
                         static_cast<const S*>(Y), count);
      TRL_HIP_LAUNCH_CHECK();
      return 0;
    }
    hipLaunchKernelGGL(k_sub<S>, dim3(grid_for(count / per + 1, kThreads, 8192)), dim3(kThreads), 0, ctx->stream, static_cast<S*>(X), static_cast<const S*>(Y),
                       count);
  });
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_panel_mult(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const double* W, void* Y, int mode)
{
  const dim3 g(grid_for(n, kThreads, 8192)), blk(kThreads);
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      S* x = static_cast<S*>(X);
      S* y = static_cast<S*>(Y);
      switch (mode) {
      case 0:
        hipLaunchKernelGGL((k_panel_mult<S, B, 0>), g, blk, 0, ctx->stream, x, W, y, n);
        break;
      case 1:
        hipLaunchKernelGGL((k_panel_mult<S, B, 1>), g, blk, 0, ctx->stream, x, W, y, n);
        break;
      case 2:
        hipLaunchKernelGGL((k_panel_mult<S, B, 2>), g, blk, 0, ctx->stream, x, W, y, n);
        break;
      case 3:
        hipLaunchKernelGGL((k_panel_mult<S, B, 3>), g, blk, 0, ctx->stream, x, W, y, n);
        break;
      default:
        hipLaunchKernelGGL((k_panel_mult<S, B, 4>), g, blk, 0, ctx->stream, x, W, y, n);
        break;
      }
    });
  });
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_small(trl_hip_ctx* ctx, int op, int b, const double* A, const double* Bm, double* C, const double* hostvals)
{
  DiagVals d{};
  if (op == 2)
    for (int i = 0; i < b; ++i) d.v[i] = hostvals[i];
  hipLaunchKernelGGL(k_small, dim3(1), dim3(256), 0, ctx->stream, op, b, A, Bm, C, d);
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_residual_sq(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const void* X, const void* Y, const void* md, const double* lambda, double* out)
{
  DiagVals lam{};
  if (lambda)
    for (int i = 0; i < b; ++i) lam.v[i] = lambda[i];
  const int nwg = int(grid_for(n, kThreads, 2 * ctx->num_cus));
  const int P = 2 * b;
  int rc = ensure_partials(ctx, size_t(nwg) * P + P);
  if (rc) return rc;
  double* part = ctx->partials;
  double* res = ctx->partials + size_t(nwg) * P;
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      hipLaunchKernelGGL((k_residual<S, B>), dim3(nwg), dim3(kThreads), 0, ctx->stream, static_cast<const S*>(X), static_cast<const S*>(Y),
                         static_cast<const S*>(md), lam, n, part);
    });
  });
  TRL_HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(unsigned(P)), dim3(64), 0, ctx->stream, part, nwg, P, res);
  TRL_HIP_LAUNCH_CHECK();
  TRL_HIP_TRY(hipMemcpyAsync(out, res, P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

} // namespace trl_hip_impl
