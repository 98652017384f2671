// Eigenvalue-adjoint SDDMM: gradient of sum_q w_q lambda_q with respect to
// the stored entries of A (and the mass diagonal of a generalized problem).
//
// d lambda_q / d A_ij = x_iq x_jq for an eigenpair with x^T M x = 1, and
// d lambda_q / d m_i = -lambda_q x_iq^2 for a diagonal M, so
//   G[p = (i, j)] = sum_q w_q x_iq x_jq       (sampled dense-dense product)
//   gm[i]         = -sum_q w_q lambda_q x_iq^2.
// One row per thread; the row's own eigenvector entries stay in registers
// per block of b vectors while the row's nonzeros gather x_j.
#include "common.hh"

using namespace trl_hip_dev;

namespace {

constexpr int kThreads = 256;

template <typename S, int B>
__global__ __launch_bounds__(kThreads) void k_sddmm(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int nblocks, const S* __restrict__ X,
                                                    const double* __restrict__ w, const double* __restrict__ lam, double* __restrict__ G,
                                                    double* __restrict__ gm, int n)
{
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t ld = size_t(n) * B;
  const int beg = rowptr[i], end = rowptr[i + 1];
  for (int p = beg; p < end; ++p) G[p] = 0.0;
  double m_acc = 0.0;
  for (int blk = 0; blk < nblocks; ++blk) {
    S xi[B];
    RowIO<S, B>::load(X + blk * ld + size_t(i) * B, xi);
    double wx[B];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      wx[c] = w[blk * B + c] * static_cast<double>(xi[c]);
      if (gm) m_acc -= lam[blk * B + c] * wx[c] * static_cast<double>(xi[c]);
    }
    for (int p = beg; p < end; ++p) {
      S xj[B];
      RowIO<S, B>::load(X + blk * ld + size_t(col[p]) * B, xj);
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < B; ++c) s = fma(wx[c], static_cast<double>(xj[c]), s);
      G[p] += s;
    }
  }
  if (gm) gm[i] = m_acc;
}

} // namespace

namespace trl_hip_impl {

int launch_sddmm(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, int b, int nblocks, const void* X, const double* w,
                 const double* lam, double* G, double* gm)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const unsigned g = unsigned((n + kThreads - 1) / kThreads);
      hipLaunchKernelGGL((k_sddmm<S, B>), dim3(g), dim3(kThreads), 0, ctx->stream, rowptr, col, nblocks, static_cast<const S*>(X), w, lam, G, gm, int(n));
      TRL_HIP_LAUNCH_CHECK();
    });
  });
  return 0;
}

} // namespace trl_hip_impl

extern "C" int trl_hip_sddmm_adjoint(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, int b, int nblocks, const void* X,
                                     const double* w, const double* lambda, double* G, double* gm)
{
  if (!ctx || !trl_hip_impl::valid_b(b) || nblocks < 1 || n > size_t(INT32_MAX)) return TRL_HIP_EINVAL;
  if (n && (!rowptr || !col || !X || !w || !G)) return TRL_HIP_EINVAL;
  if (gm && !lambda) return TRL_HIP_EINVAL;
  if (!n) return 0;
  return trl_hip_impl::launch_sddmm(ctx, t, n, rowptr, col, b, nblocks, X, w, lambda, G, gm);
}
