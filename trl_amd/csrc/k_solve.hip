// Inner solver of the shift-invert operator: K X = R for b right-hand sides
// at once (K = A - sigma B symmetric positive definite), by the conjugate
// gradient method with the Jacobi preconditioner (Hestenes & Stiefel 1952;
// the preconditioned form of Golub & Van Loan, Matrix Computations, 4th ed.,
// Alg. 11.5.1), one independent recurrence per column, all columns advanced
// by the same launches.
//
// One iteration is four launches on the context stream and no host
// synchronisation:
//   Q = K P                      the operator kernel (any trl_hip_spmat form)
//   k_pcg_pq                     per-workgroup partial sums of P .* Q
//   k_pcg_update                 alpha = rz / p'q (every workgroup reduces the
//                                partials itself, in the same order, so all
//                                agree bit for bit); X += alpha P,
//                                R -= alpha Q; partial sums of R .* D^-1 R and
//                                R .* R
//   k_pcg_direction              beta = rz_new / rz; P = D^-1 R + beta P;
//                                workgroup 0 records rz_new and |R|^2
// The scalars live in a small device state: rz of the current and the next
// iteration (two slots, chosen by the iteration's parity, so that no
// workgroup overwrites a value another still reads), |R|^2, |R_0|^2 and a
// flag for a direction with p'Kp <= 0 (K not positive definite).  The host
// reads the state every few iterations to stop.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hh"

using namespace trl_hip_dev;

namespace {

constexpr int kThreads = 256;
constexpr int kMaxB = 16;
// workgroups of the reduction kernels (grid-stride over the rows)
constexpr int kMaxPartialWgs = 512;

// state layout (doubles): rz[2][kMaxB], rr[kMaxB], r0[kMaxB], flag
constexpr int kRz = 0, kRr = 2 * kMaxB, kR0 = 3 * kMaxB, kFlag = 4 * kMaxB;

// Sums of `nsum` interleaved per-workgroup partials (partials[w * nsum + c],
// w < G) into out[c], one wave per column, lanes over workgroups in a fixed
// order: the same bits in every workgroup that calls it.
__device__ inline void reduce_partials(const double* __restrict__ partials, int G, int nsum, double* out)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int c = wave; c < nsum; c += nw) {
    double a = 0.0;
    for (int w = lane; w < G; w += 64) a += partials[size_t(w) * nsum + c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) out[c] = a;
  }
  __syncthreads();
}

// Workgroup sums of per-thread values v[0..nsum) into partials[blockIdx.x * nsum + c].
template <int NS>
__device__ inline void block_partials(const double (&v)[NS], int nsum, double* __restrict__ partials)
{
  __shared__ double red[kThreads / 64][2 * kMaxB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NS; ++c) {
    double x = v[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0 && c < nsum) red[wave][c] = x;
  }
  __syncthreads();
  if (threadIdx.x < unsigned(nsum)) {
    double s = 0.0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w][threadIdx.x];
    partials[size_t(blockIdx.x) * nsum + threadIdx.x] = s;
  }
}

// X = 0, R = Rhs, P = D^-1 Rhs; partials of Rhs .* D^-1 Rhs and Rhs .* Rhs.
template <int B>
__global__ __launch_bounds__(kThreads) void k_pcg_init(const double* __restrict__ rhs, const double* __restrict__ dinv, size_t n, double* __restrict__ X,
                                                       double* __restrict__ R, double* __restrict__ P, double* __restrict__ partials)
{
  double v[2 * B];
#pragma unroll
  for (int c = 0; c < 2 * B; ++c) v[c] = 0.0;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const double d = dinv[i];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      const double r = rhs[i * B + c];
      X[i * B + c] = 0.0;
      R[i * B + c] = r;
      P[i * B + c] = d * r;
      v[c] = fma(r, d * r, v[c]);
      v[B + c] = fma(r, r, v[B + c]);
    }
  }
  block_partials<2 * B>(v, 2 * B, partials);
}

// state: rz[slot 0] = sum Rhs .* D^-1 Rhs, rr = r0 = |Rhs|^2, flag = 0.
template <int B>
__global__ __launch_bounds__(kThreads) void k_pcg_init_finish(const double* __restrict__ partials, int G, double* __restrict__ state)
{
  __shared__ double s[2 * kMaxB];
  reduce_partials(partials, G, 2 * B, s);
  if (threadIdx.x < unsigned(B)) {
    state[kRz + threadIdx.x] = s[threadIdx.x];
    state[kRr + threadIdx.x] = s[B + threadIdx.x];
    state[kR0 + threadIdx.x] = s[B + threadIdx.x];
  }
  if (threadIdx.x == 0) state[kFlag] = 0.0;
}

template <int B>
__global__ __launch_bounds__(kThreads) void k_pcg_pq(const double* __restrict__ P, const double* __restrict__ Q, size_t n, double* __restrict__ partials)
{
  double v[B];
#pragma unroll
  for (int c = 0; c < B; ++c) v[c] = 0.0;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
#pragma unroll
    for (int c = 0; c < B; ++c) v[c] = fma(P[i * B + c], Q[i * B + c], v[c]);
  block_partials<B>(v, B, partials);
}

// alpha = rz / p'q per column (0 where rz = 0: a column already solved
// exactly); p'q <= 0 with rz > 0 flags an indefinite K and takes alpha = 0.
template <int B>
__global__ __launch_bounds__(kThreads) void k_pcg_update(const double* __restrict__ P, const double* __restrict__ Q, const double* __restrict__ dinv, size_t n,
                                                         int G, const double* __restrict__ pq_partials, double* __restrict__ state, int slot,
                                                         double* __restrict__ X, double* __restrict__ R, double* __restrict__ partials)
{
  __shared__ double pq[kMaxB], alpha[kMaxB];
  reduce_partials(pq_partials, G, B, pq);
  if (threadIdx.x < unsigned(B)) {
    const double rz = state[kRz + slot * kMaxB + threadIdx.x], d = pq[threadIdx.x];
    double a = 0.0;
    if (rz > 0.0) {
      if (d > 0.0)
        a = rz / d;
      else if (blockIdx.x == 0)
        state[kFlag] = 1.0;
    }
    alpha[threadIdx.x] = a;
  }
  __syncthreads();
  double v[2 * B];
#pragma unroll
  for (int c = 0; c < 2 * B; ++c) v[c] = 0.0;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const double d = dinv[i];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      const double a = alpha[c];
      X[i * B + c] = fma(a, P[i * B + c], X[i * B + c]);
      const double r = fma(-a, Q[i * B + c], R[i * B + c]);
      R[i * B + c] = r;
      v[c] = fma(r, d * r, v[c]);
      v[B + c] = fma(r, r, v[B + c]);
    }
  }
  block_partials<2 * B>(v, 2 * B, partials);
}

// beta = rz_new / rz; P = D^-1 R + beta P; workgroup 0 stores rz_new in the
// other slot and |R|^2.
template <int B>
__global__ __launch_bounds__(kThreads) void k_pcg_direction(const double* __restrict__ R, const double* __restrict__ dinv, size_t n, int G,
                                                            const double* __restrict__ rz_partials, double* __restrict__ state, int slot,
                                                            double* __restrict__ P)
{
  __shared__ double s[2 * kMaxB], beta[kMaxB];
  reduce_partials(rz_partials, G, 2 * B, s);
  if (threadIdx.x < unsigned(B)) {
    const double rz = state[kRz + slot * kMaxB + threadIdx.x];
    beta[threadIdx.x] = rz > 0.0 ? s[threadIdx.x] / rz : 0.0;
  }
  __syncthreads();
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const double d = dinv[i];
#pragma unroll
    for (int c = 0; c < B; ++c) P[i * B + c] = fma(beta[c], P[i * B + c], d * R[i * B + c]);
  }
  if (blockIdx.x == 0 && threadIdx.x < unsigned(B)) {
    state[kRz + (1 - slot) * kMaxB + threadIdx.x] = s[threadIdx.x];
    state[kRr + threadIdx.x] = s[B + threadIdx.x];
  }
}

} // namespace

namespace trl_hip_impl {

size_t pcg_workspace_bytes(size_t n, int b)
{
  // three panels (R, P, Q), the partials of the p'q sums (kMaxPartialWgs x
  // kMaxB) and of the rz / |r|^2 sums (kMaxPartialWgs x 2 kMaxB), the state
  return (3 * n * size_t(b) + size_t(kMaxPartialWgs) * 3 * kMaxB + 8 * kMaxB) * sizeof(double);
}

int launch_pcg(trl_hip_ctx* ctx, size_t n, int b, const trl_hip_spmat& K, const double* dinv, const double* rhs, double* X, double tol, int maxit,
               void* work, int* iters, double* relres)
{
  if (b < 1 || b > kMaxB || !valid_b(b)) return TRL_HIP_EINVAL;
  const size_t P = n * size_t(b);
  double* R = static_cast<double*>(work);
  double* Pd = R + P;
  double* Q = Pd + P;
  double* pq_part = Q + P;                                  // kMaxPartialWgs * kMaxB
  double* rz_part = pq_part + size_t(kMaxPartialWgs) * kMaxB; // kMaxPartialWgs * 2 kMaxB
  double* state = rz_part + size_t(kMaxPartialWgs) * 2 * kMaxB;
  const int G = int(std::min<size_t>((n + kThreads - 1) / kThreads, kMaxPartialWgs));
  const dim3 grid{unsigned(G)}, blk{unsigned(kThreads)};
  std::vector<double> h(8 * kMaxB);
  int rc = 0;
  int it = 0;
  double worst = 0.0;
  // the host checks the state every `check` iterations (one stream sync)
  const int check = 8;
  TRL_DISPATCH_B(b, {
    hipLaunchKernelGGL((k_pcg_init<B>), grid, blk, 0, ctx->stream, rhs, dinv, n, X, R, Pd, rz_part);
    TRL_HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_pcg_init_finish<B>), dim3(1), blk, 0, ctx->stream, rz_part, G, state);
    TRL_HIP_LAUNCH_CHECK();
    for (;;) {
      // convergence: |R|^2 <= tol^2 |R_0|^2 in every column
      if (it % check == 0 || it >= maxit) {
        if (hipMemcpyAsync(h.data(), state, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
          return TRL_HIP_EINVAL;
        if (h[kFlag] != 0.0) {
          rc = TRL_HIP_EINDEF;
          break;
        }
        bool done = true;
        worst = 0.0;
        for (int c = 0; c < B; ++c) {
          const double r0 = h[kR0 + c], rr = h[kRr + c];
          const double rel = r0 > 0.0 ? std::sqrt(rr / r0) : 0.0;
          worst = std::max(worst, rel);
          if (!(rel <= tol)) done = false;
        }
        if (done) break;
        if (it >= maxit) {
          rc = TRL_HIP_ENOCONV;
          break;
        }
      }
      const int slot = it & 1;
      rc = launch_spmm(ctx, TRL_HIP_F64, n, B, K, Pd, Q, nullptr, false, 0.0, 0.0, 0.0);
      if (rc) break;
      hipLaunchKernelGGL((k_pcg_pq<B>), grid, blk, 0, ctx->stream, Pd, Q, n, pq_part);
      TRL_HIP_LAUNCH_CHECK();
      hipLaunchKernelGGL((k_pcg_update<B>), grid, blk, 0, ctx->stream, Pd, Q, dinv, n, G, pq_part, state, slot, X, R, rz_part);
      TRL_HIP_LAUNCH_CHECK();
      hipLaunchKernelGGL((k_pcg_direction<B>), grid, blk, 0, ctx->stream, R, dinv, n, G, rz_part, state, slot, Pd);
      TRL_HIP_LAUNCH_CHECK();
      ++it;
    }
  });
  if (iters) *iters = it;
  if (relres) *relres = worst;
  return rc;
}

} // namespace trl_hip_impl
