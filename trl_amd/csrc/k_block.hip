// Tall-skinny block products of the block-Lanczos hot loop:
//   Gram     C = V[0:k]^T diag(m) W            (MFMA, two-level deterministic reduction)
//   update   W -= V[0:k] C                      (streaming, C in LDS)
//   CholQR   W <- W R^{-1}, R^T R = W^T M W     (Gram + FP64 Cholesky on the device + scale)
//   rotate   Out_j = sum_i V_i Y(i, j)          (MFMA, Y in LDS)
//
// MFMA operand layout used throughout (v_mfma_{f32,f64}_16x16x4): lane l
// supplies A[i = l % 16][k = l / 16] and B[k = l / 16][j = l % 16].  The
// 16x16 result D[i][j] lives in lane j + 16*q, element e, with
//   FP32: i = 4*q + e   (four consecutive rows per lane)
//   FP64: i = 4*e + q   (rows strided by four)
// (the f32/f64 output groupings of the 16x16x4 instructions differ).
#include <cstdlib>
#include <string>

#include <mutex>
#include <type_traits>

#include "common.hh"

using namespace trl_hip_dev;

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename S>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = v4d;
  __device__ static inline acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct Mfma<float> {
  using acc_t = v4f;
  __device__ static inline acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

constexpr int pad_b(int b) { return b <= 1 ? 1 : b <= 2 ? 2 : b <= 4 ? 4 : b <= 8 ? 8 : 16; }

// Row of the 16x16 MFMA result held by element e of a lane in row group q.
template <typename S>
__device__ constexpr int mfma_row(int q, int e)
{
  return sizeof(S) == 8 ? 4 * e + q : 4 * q + e;
}

// Quad-lane exchanges (DPP quad permutes, no memory): the value of lane
// (lane ^ 1) or (lane ^ 2) within each group of four lanes.
template <int CTRL>
__device__ inline float dpp_f(float x)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline double dpp_f(double x)
{
  const long long u = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_mov_dpp(int(u & 0xffffffff), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(int(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
constexpr int kDppXor1 = 0xB1; // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E; // quad_perm [2, 3, 0, 1]

// 2 x 2 transposes of registers (e0, e0 + D) across lanes (lane, lane ^ L)
// in place: afterwards the lane with bit L clear holds what lane and its
// partner held in register e0, the other what they held in e0 + D.
template <int CTRL, int L, int D, typename A>
__device__ inline void pair_transpose(A& r, int e0, int lane)
{
  const bool hi = (lane & L) != 0;
  const auto send = hi ? r[e0] : r[e0 + D];
  const auto recv = dpp_f<CTRL>(send);
  if (hi)
    r[e0] = recv;
  else
    r[e0 + D] = recv;
}

constexpr int kThreads = 256;
// Gram: row quads per load group (16 rows)
constexpr int kU = 4;
// Gram workgroups per compute unit
constexpr int kGramWgPerCu = 4;

// ---------------------------------------------------------------------------
// Gram.  A tile stacks 16/BP blocks of V (columns padded to BP = pow2(b));
// B is diag(m) W padded to 16 columns; K runs over rows, 4 per instruction.
// A workgroup has 4 waves: TG tile groups x RS row subsets.  Wave (tg, rs)
// handles tiles t = tg + TG*q (q < TPW) on the row quads of subset rs and
// writes its partial sums to partials[(blockIdx.x*RS + rs)*P + idx], P = k*b*b.
template <typename S, int B, int TPW>
__global__ __launch_bounds__(kThreads) void k_gram(const S* __restrict__ V, size_t ldv, int k, const S* __restrict__ W, const S* __restrict__ md, size_t n,
                                                   int TG, double* __restrict__ partials, const int* __restrict__ gate)
{
  if (gate && *gate == 0) return;
  using M = Mfma<S>;
  using acc_t = typename M::acc_t;
  constexpr int BP = pad_b(B), BPT = 16 / BP;
  const int ntile = (k + BPT - 1) / BPT;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int RS = 4 / TG;
  const int tg = wave % TG, rs = wave / TG;
  const int i16 = lane & 15, q4 = lane >> 4;
  const int bit = i16 / BP, col = i16 % BP;

  size_t aoff[TPW];
  bool aval[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = tg + TG * q;
    const int j = t * BPT + bit;
    aval[q] = (t < ntile) && (j < k) && (col < B);
    aoff[q] = aval[q] ? size_t(j) * ldv + col : 0;
  }
  const bool wval = i16 < B;

  // Each row wave walks groups of kU consecutive row quads (4*kU rows); all
  // loads of a group are issued before its MFMAs so that enough bytes are in
  // flight to cover HBM latency.
  const size_t nquads = (n + 3) / 4;
  const size_t ngroups = (nquads + kU - 1) / kU;
  const size_t gstride = size_t(gridDim.x) * RS;

  // FP64: one accumulator; FP32: the MFMA accumulates one group (16 rows) in
  // FP32 and is folded into an error-free (hi, lo) pair by TwoSum
  acc_t acc[TPW];
  acc_t lo[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    acc[q] = acc_t{0, 0, 0, 0};
    lo[q] = acc_t{0, 0, 0, 0};
  }

  for (size_t g = size_t(blockIdx.x) * RS + rs; g < ngroups; g += gstride) {
    S bw[kU];
    S a[kU][TPW];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const size_t r = (g * kU + u) * 4 + q4;
      const bool rok = r < n;
      bw[u] = 0;
      if (rok && wval) {
        bw[u] = W[r * B + i16];
        if (md) bw[u] *= md[r];
      }
#pragma unroll
      for (int q = 0; q < TPW; ++q) a[u][q] = (rok && aval[q]) ? V[aoff[q] + r * B] : S(0);
    }
    if constexpr (sizeof(S) == 8) {
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int q = 0; q < TPW; ++q) acc[q] = M::op(a[u][q], bw[u], acc[q]);
    }
    else {
      acc_t cur[TPW];
#pragma unroll
      for (int q = 0; q < TPW; ++q) cur[q] = acc_t{0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int q = 0; q < TPW; ++q) cur[q] = M::op(a[u][q], bw[u], cur[q]);
#pragma unroll
      for (int q = 0; q < TPW; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float s, err;
          two_sum(acc[q][e], cur[q][e], s, err);
          acc[q][e] = s;
          lo[q][e] += err;
        }
    }
  }

  // entry-major partials: entry idx of (workgroup, row subset) w at idx*NW + w
  const size_t NW = size_t(gridDim.x) * RS;
  double* out = partials + size_t(blockIdx.x) * RS + rs;
  if (!wval) return;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = tg + TG * q;
    if (t >= ntile) continue;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = mfma_row<S>(q4, e);
      const int blk = t * BPT + i / BP, ci = i % BP;
      if (blk < k && ci < B) {
        double v = static_cast<double>(acc[q][e]);
        if constexpr (sizeof(S) == 4) v += static_cast<double>(lo[q][e]);
        out[((size_t(blk) * B + ci) * B + i16) * NW] = v;
      }
    }
  }
}

// The Gram through range-checked buffer loads, row groups in two register
// sets: the loads of the next group are issued (kept ahead by a scheduling
// barrier) before the products of the current one.  Rows past n, padding
// columns and tiles past k read zeros through offsets past their resource
// (the W and md panels, and each tile's BPT blocks of V, below 2^31 bytes:
// checked by the launcher), so no lane branches around a load and groups
// past the last one add exact zeros.  Same tiles, row groups, accumulation
// order and partial-sum layout as k_gram.
template <typename S, int B, int TPW>
__global__ __launch_bounds__(kThreads) void k_gram_buf(const S* __restrict__ V, size_t ldv, int k, const S* __restrict__ W, const S* __restrict__ md,
                                                       size_t n, int TG, double* __restrict__ partials, const int* __restrict__ gate)
{
  if (gate && *gate == 0) return;
  using M = Mfma<S>;
  using acc_t = typename M::acc_t;
  constexpr int BP = pad_b(B), BPT = 16 / BP;
  constexpr unsigned kBad = 1u << 31;
  const int ntile = (k + BPT - 1) / BPT;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int RS = 4 / TG;
  const int tg = wave % TG, rs = wave / TG;
  const int i16 = lane & 15, q4 = lane >> 4;
  const int bit = i16 / BP, col = i16 % BP;
  const unsigned ldv32 = unsigned(ldv);

  __amdgpu_buffer_rsrc_t vr[TPW];
  unsigned aofs[TPW], abad[TPW];
  bool aval[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = tg + TG * q;
    const int j0 = t * BPT, j = j0 + bit;
    aval[q] = (t < ntile) && (j < k) && (col < B);
    const int span = t < ntile ? ((k - j0) < BPT ? (k - j0) : BPT) : 0;
    vr[q] = buf_rsrc_uniform(V + size_t(t < ntile ? j0 : 0) * ldv, size_t(span) * ldv * sizeof(S));
    aofs[q] = (unsigned(bit) * ldv32 + unsigned(col)) * unsigned(sizeof(S));
    abad[q] = aval[q] ? 0u : kBad;
  }
  const bool wval = i16 < B;
  const __amdgpu_buffer_rsrc_t wr = buf_rsrc(W, n * B * sizeof(S));
  const __amdgpu_buffer_rsrc_t mr = buf_rsrc(md, md ? n * sizeof(S) : 0);
  const unsigned wofs = unsigned(i16) * unsigned(sizeof(S)), wbad = wval ? 0u : kBad;

  const size_t nquads = (n + 3) / 4;
  const size_t ngroups = (nquads + kU - 1) / kU;
  const size_t gstride = size_t(gridDim.x) * RS;

  acc_t acc[TPW];
  acc_t lo[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    acc[q] = acc_t{0, 0, 0, 0};
    lo[q] = acc_t{0, 0, 0, 0};
  }
  // the W and V values of row group g held by this lane (zeros past n)
  // (the md factor is applied where the products use it, so that no load
  // waits before the next group's loads are issued)
  auto load_group = [&](size_t g, S (&bw)[kU], S (&mv)[kU], S (&a)[kU][TPW]) {
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const size_t r = (g * kU + u) * 4 + q4;
      const unsigned rbad = unsigned(r >= n) << 31;
      const unsigned rofs = unsigned(r) * unsigned(B * sizeof(S));
      bw[u] = buf_load<S>(wr, (rofs + wofs) | wbad | rbad);
      mv[u] = buf_load<S>(mr, (unsigned(r) * unsigned(sizeof(S))) | rbad);
#pragma unroll
      for (int q = 0; q < TPW; ++q) a[u][q] = buf_load<S>(vr[q], (rofs + aofs[q]) | abad[q] | rbad);
    }
  };
  auto mma_group = [&](const S (&bw0)[kU], const S (&mv)[kU], const S (&a)[kU][TPW]) {
    S bw[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) bw[u] = md ? bw0[u] * mv[u] : bw0[u];
    if constexpr (sizeof(S) == 8) {
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int q = 0; q < TPW; ++q) acc[q] = M::op(a[u][q], bw[u], acc[q]);
    }
    else {
      acc_t cur[TPW];
#pragma unroll
      for (int q = 0; q < TPW; ++q) cur[q] = acc_t{0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int q = 0; q < TPW; ++q) cur[q] = M::op(a[u][q], bw[u], cur[q]);
#pragma unroll
      for (int q = 0; q < TPW; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float s, err;
          two_sum(acc[q][e], cur[q][e], s, err);
          acc[q][e] = s;
          lo[q][e] += err;
        }
    }
  };
  size_t g = size_t(blockIdx.x) * RS + rs;
  S bwA[kU], mA[kU], aA[kU][TPW], bwB[kU], mB[kU], aB[kU][TPW];
  load_group(g, bwA, mA, aA);
  for (; g < ngroups; g += 2 * gstride) {
    load_group(g + gstride, bwB, mB, aB);
    __builtin_amdgcn_sched_barrier(0);
    mma_group(bwA, mA, aA);
    load_group(g + 2 * gstride, bwA, mA, aA);
    __builtin_amdgcn_sched_barrier(0);
    mma_group(bwB, mB, aB);
  }

  const size_t NW = size_t(gridDim.x) * RS;
  double* out = partials + size_t(blockIdx.x) * RS + rs;
  if (!wval) return;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = tg + TG * q;
    if (t >= ntile) continue;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = mfma_row<S>(q4, e);
      const int blk = t * BPT + i / BP, ci = i % BP;
      if (blk < k && ci < B) {
        double v = static_cast<double>(acc[q][e]);
        if constexpr (sizeof(S) == 4) v += static_cast<double>(lo[q][e]);
        out[((size_t(blk) * B + ci) * B + i16) * NW] = v;
      }
    }
  }
}

// out[p] = sum_{w < nw} partials[p*nw + w] (entry-major), summed in a fixed
// order (deterministic): one workgroup per entry, each lane strides over w
// with four accumulators, then a butterfly per wave and the waves in order.
__global__ __launch_bounds__(kThreads) void k_reduce(const double* __restrict__ partials, int nw, int P, double* __restrict__ out,
                                                     const int* __restrict__ gate)
{
  if (gate && *gate == 0) return;
  __shared__ double red[kThreads / 64];
  const int p = blockIdx.x;
  if (p >= P) return;
  const double* src = partials + size_t(p) * nw;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int w = threadIdx.x;
  for (; w + 3 * kThreads < nw; w += 4 * kThreads) {
    a0 += src[w];
    a1 += src[w + kThreads];
    a2 += src[w + 2 * kThreads];
    a3 += src[w + 3 * kThreads];
  }
  for (; w < nw; w += kThreads) a0 += src[w];
  double v = (a0 + a1) + (a2 + a3);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) t += red[q];
    out[p] = t;
  }
}

// Factor G = R^T R (G symmetric B x B, B <= 8, its lower triangle read as
// an LLT reads it), set Rout = R Rprev (or R) and Rinv = R^{-1}, one thread
// per column: row k of R is R[k][j] = (G[j][k] - sum_{p<k} R[p][k] R[p][j]) /
// R[k][k] (the terms subtracted in p order) for all j >= k at once, R kept
// in LDS (Rs, B x B) between the steps, T (B) the step's partial sums;
// thread j then forms column j of R^{-1} by back substitution and of
// R Rprev.  Every thread of the workgroup must call it (it holds workgroup
// barriers); threads j >= B only take part in them.  On breakdown (a pivot
// that is not positive) the first failing 1-based pivot goes to flags[0] if
// it is clear, and Rout and Rinv are set to the identity.
template <int B>
__device__ void chol_lanes(const double* G, const double* __restrict__ Rprev, double* __restrict__ Rout, double* __restrict__ Rinv, int* __restrict__ flags,
                           double* Rs, double* T)
{
  static_assert(B <= 8, "one thread per column");
  const int j = threadIdx.x;
  const bool mine = j < B;
  int fail = 0;
#pragma unroll
  for (int k = 0; k < B; ++k) {
    if (mine && j >= k) {
      double t = G[j * B + k];
#pragma unroll
      for (int p = 0; p < k; ++p) t -= Rs[p * B + k] * Rs[p * B + j];
      T[j] = t;
    }
    __syncthreads();
    const double d = T[k];
    if (!(d > 0.0) && !fail) fail = k + 1;
    const double rkk = sqrt(d > 0.0 ? d : 1.0);
    if (mine) Rs[k * B + j] = j == k ? rkk : (j > k ? T[j] / rkk : 0.0);
    __syncthreads();
  }
  if (!mine) return;
  if (fail) {
    if (j == 0 && flags[0] == 0) flags[0] = fail;
#pragma unroll
    for (int i = 0; i < B; ++i) Rout[i * B + j] = Rinv[i * B + j] = (i == j) ? 1.0 : 0.0;
    return;
  }
  double xcol[B];
#pragma unroll
  for (int i = B - 1; i >= 0; --i) {
    double t = 0.0;
#pragma unroll
    for (int p = i + 1; p < B; ++p) t += Rs[i * B + p] * xcol[p]; // X[p][j] = 0 for p > j
    xcol[i] = i > j ? 0.0 : (i == j ? 1.0 / Rs[i * B + i] : -t / Rs[i * B + i]);
  }
#pragma unroll
  for (int i = 0; i < B; ++i) {
    double v = Rs[i * B + j];
    if (Rprev) {
      v = 0.0;
#pragma unroll
      for (int p = i; p < B; ++p) v += Rs[i * B + p] * Rprev[p * B + j];
    }
    Rinv[i * B + j] = xcol[i];
    Rout[i * B + j] = i > j ? 0.0 : v;
  }
}

// Factor G = R^T R (G symmetric b x b in LDS, lower triangle read as an LLT
// does), set Rout = R Rprev (or R) and Rinv = R^{-1}; on breakdown record the
// 1-based pivot in the sticky flag and return identities.  One thread.
__device__ void chol_bxb(const double* G, int b, const double* Rprev, double* Rout, double* Rinv, int* flags, double (*R)[16], double (*X)[16])
{
  int fail = 0;
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j) R[i][j] = X[i][j] = 0.0;
  for (int j = 0; j < b && !fail; ++j) {
    double d = G[j * b + j];
    for (int p = 0; p < j; ++p) d -= R[p][j] * R[p][j];
    if (!(d > 0.0)) {
      fail = j + 1;
      break;
    }
    const double rjj = sqrt(d);
    R[j][j] = rjj;
    for (int i = j + 1; i < b; ++i) {
      double s = G[i * b + j];
      for (int p = 0; p < j; ++p) s -= R[p][j] * R[p][i];
      R[j][i] = s / rjj;
    }
  }
  if (fail) {
    if (flags[0] == 0) flags[0] = fail;
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < b; ++j) {
        Rout[i * b + j] = (i == j) ? 1.0 : 0.0;
        Rinv[i * b + j] = (i == j) ? 1.0 : 0.0;
      }
    return;
  }
  for (int j = 0; j < b; ++j) {
    X[j][j] = 1.0 / R[j][j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int p = i + 1; p <= j; ++p) s += R[i][p] * X[p][j];
      X[i][j] = -s / R[i][i];
    }
  }
  for (int i = 0; i < b; ++i)
    for (int j = 0; j < b; ++j) Rinv[i * b + j] = X[i][j];
  for (int i = 0; i < b; ++i)
    for (int j = 0; j < b; ++j) {
      double v = 0.0;
      if (j >= i) {
        if (Rprev)
          for (int p = i; p <= j; ++p) v += R[i][p] * Rprev[p * b + j];
        else
          v = R[i][j];
      }
      Rout[i * b + j] = v;
    }
}

// One Cholesky-QR pass over the panel W (n x B), one row per thread:
//   SCALE: W <- W Rinv_prev first (in place, FP64 arithmetic);
//   G = W^T diag(m) W accumulated in FP64 over the stored values (upper
//   triangle, B(B+1)/2 entries), reduced per workgroup (shuffles, then LDS);
//   the last workgroup to finish sums the workgroup partials in fixed order
//   and factors G on the device (chol_bxb), then re-arms the counter.
// The end of a one-pass Gram + Cholesky kernel.  g: this thread's packed
// upper triangle of the Gram contributions (NP = B(B+1)/2 entries).  The
// workgroup sums them (butterfly inside each wave, then the waves in order)
// into partials, stored entry-major (entry i of workgroup w at i*G + w); the
// workgroup that takes the last ticket on *counter then sums every
// workgroup's partials in a fixed order (thread t: workgroups t, t + 256,
// ...; then the same butterfly and wave order: deterministic for a given
// grid), factors G = R^T R in FP64 (R times Rprev into Rout when Rprev is
// given, this pass's R^{-1} into Rinv_out, a breakdown into flags) and
// re-arms the counter.  red, Gs, Rl, Xl and ticket are the caller's LDS.
template <int B>
__device__ void finish_gram_chol(double (&g)[B * (B + 1) / 2], double* __restrict__ partials, unsigned* __restrict__ counter,
                                 const double* __restrict__ Rprev, double* __restrict__ Rout, double* __restrict__ Rinv_out, int* __restrict__ flags,
                                 double (*red)[B * (B + 1) / 2], double* Gs, double (*Rl)[16], double (*Xl)[16], unsigned& ticket)
{
  constexpr int NP = B * (B + 1) / 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned G = gridDim.x;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    double v = g[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NP; i += blockDim.x) {
    double t = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) t += red[q][i];
    partials[size_t(i) * G + blockIdx.x] = t;
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) ticket = atomicAdd(counter, 1u);
  __syncthreads();
  if (ticket != G - 1) return;
  __threadfence();
  // last workgroup: every thread sums the workgroups w = t, t + 256, ... of
  // all entries (independent loads), then the same butterfly and wave order
  // as above (fixed order: deterministic for a given grid)
#pragma unroll
  for (int i = 0; i < NP; ++i) g[i] = 0.0;
  for (unsigned w = threadIdx.x; w < G; w += kThreads)
#pragma unroll
    for (int i = 0; i < NP; ++i) g[i] += partials[size_t(i) * G + w];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    double v = g[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NP; i += blockDim.x) {
    double t = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) t += red[q][i];
    // unpack the upper-triangle index i into (r, c) and mirror
    int r = 0, rem = i;
    while (rem >= B - r) {
      rem -= B - r;
      ++r;
    }
    const int c = r + rem;
    Gs[r * B + c] = t;
    Gs[c * B + r] = t;
  }
  __syncthreads();
  if constexpr (B <= 8)
    chol_lanes<B>(Gs, Rprev, Rout, Rinv_out, flags, &Rl[0][0], &Xl[0][0]);
  else if (threadIdx.x == 0)
    chol_bxb(Gs, B, Rprev, Rout, Rinv_out, flags, Rl, Xl);
  if (threadIdx.x == 0) *counter = 0u;
}

template <typename S, int B, bool SCALE>
__global__ __launch_bounds__(kThreads) void k_cholqr_pass(S* __restrict__ W, const S* __restrict__ md, size_t n, const double* __restrict__ Rinv_prev,
                                                          double* __restrict__ partials, unsigned* __restrict__ counter, const double* __restrict__ Rprev,
                                                          double* __restrict__ Rout, double* __restrict__ Rinv_out, int* __restrict__ flags)
{
  constexpr int NP = B * (B + 1) / 2;
  __shared__ double Rs[B * B];
  __shared__ double red[kThreads / 64][NP];
  __shared__ double Gs[B * B];
  __shared__ double Rl[16][16], Xl[16][16];
  __shared__ unsigned ticket;
  if constexpr (SCALE) {
    for (int i = threadIdx.x; i < B * B; i += blockDim.x) Rs[i] = Rinv_prev[i];
    __syncthreads();
  }
  double g[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) g[i] = 0.0;
  // two rows per iteration: both loads are in flight before the arithmetic
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t row = size_t(blockIdx.x) * blockDim.x + threadIdx.x; row < n; row += 2 * stride) {
    const size_t rows[2] = {row, row + stride};
    const bool has2 = rows[1] < n;
    S w[2][B];
    RowIO<S, B>::load(W + rows[0] * B, w[0]);
    if (has2) RowIO<S, B>::load(W + rows[1] * B, w[1]);
    double mm[2] = {1.0, 1.0};
    if (md) {
      mm[0] = static_cast<double>(md[rows[0]]);
      if (has2) mm[1] = static_cast<double>(md[rows[1]]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (u == 1 && !has2) break;
      double wd[B];
      if constexpr (SCALE) {
        double acc[B];
#pragma unroll
        for (int c = 0; c < B; ++c) acc[c] = 0.0;
#pragma unroll
        for (int cp = 0; cp < B; ++cp) {
          const double x = static_cast<double>(w[u][cp]);
#pragma unroll
          for (int c = cp; c < B; ++c) acc[c] = fma(x, Rs[cp * B + c], acc[c]);
        }
#pragma unroll
        for (int c = 0; c < B; ++c) {
          w[u][c] = static_cast<S>(acc[c]);
          wd[c] = static_cast<double>(w[u][c]);
        }
        RowIO<S, B>::store(W + rows[u] * B, w[u]);
      }
      else {
#pragma unroll
        for (int c = 0; c < B; ++c) wd[c] = static_cast<double>(w[u][c]);
      }
      int idx = 0;
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const double wi = wd[i] * mm[u];
#pragma unroll
        for (int j = i; j < B; ++j) {
          g[idx] = fma(wi, wd[j], g[idx]);
          ++idx;
        }
      }
    }
  }
  finish_gram_chol<B>(g, partials, counter, Rprev, Rout, Rinv_out, flags, red, Gs, Rl, Xl, ticket);
}

// W <- W Rinv (Rinv upper triangular), applied in FP64.
template <typename S, int B>
__global__ __launch_bounds__(kThreads) void k_scale(S* __restrict__ W, const double* __restrict__ Rinv, size_t n)
{
  __shared__ double Rs[B * B];
  for (int i = threadIdx.x; i < B * B; i += blockDim.x) Rs[i] = Rinv[i];
  __syncthreads();
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t row = size_t(blockIdx.x) * blockDim.x + threadIdx.x; row < n; row += stride) {
    S in[B];
    RowIO<S, B>::load(W + row * B, in);
    double acc[B];
#pragma unroll
    for (int c = 0; c < B; ++c) acc[c] = 0.0;
#pragma unroll
    for (int cp = 0; cp < B; ++cp) {
      const double x = static_cast<double>(in[cp]);
#pragma unroll
      for (int c = cp; c < B; ++c) acc[c] = fma(x, Rs[cp * B + c], acc[c]);
    }
    S out[B];
#pragma unroll
    for (int c = 0; c < B; ++c) out[c] = static_cast<S>(acc[c]);
    RowIO<S, B>::store(W + row * B, out);
  }
}

// W -= sum_{j<k} V_j C_j ; C (k*B x B, FP64) staged in LDS in the panel type.
template <typename S, int B>
__global__ __launch_bounds__(kThreads) void k_update(const S* __restrict__ V, size_t ldv, int k, const double* __restrict__ C, S* __restrict__ W, size_t n,
                                                     const int* __restrict__ gate)
{
  if (gate && *gate == 0) return;
  extern __shared__ unsigned char smem_raw[];
  S* Cs = reinterpret_cast<S*>(smem_raw);
  const int P = k * B * B;
  for (int i = threadIdx.x; i < P; i += blockDim.x) Cs[i] = static_cast<S>(C[i]);
  __syncthreads();
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t row = size_t(blockIdx.x) * blockDim.x + threadIdx.x; row < n; row += stride) {
    S acc[B];
    RowIO<S, B>::load(W + row * B, acc);
    for (int j = 0; j < k; ++j) {
      S v[B];
      RowIO<S, B>::load(V + size_t(j) * ldv + row * B, v);
      const S* Cj = Cs + j * B * B;
#pragma unroll
      for (int cp = 0; cp < B; ++cp)
#pragma unroll
        for (int c = 0; c < B; ++c) acc[c] = fma(-v[cp], Cj[cp * B + c], acc[c]);
    }
    RowIO<S, B>::store(W + row * B, acc);
  }
}

// Out_j = sum_{i<m} V_i Y(i, jb0 + j) for output blocks jb0 <= j < jb0 + nb.
// The needed part of Y (m*B x nb*B, padded to NT*16 columns) is staged in
// LDS in the panel type; each wave computes 16 rows x NT*16 columns with K
// running over the m*B basis columns.
// Rotation operand loads: VEC consecutive columns of a basis row per lane
// and load (16 bytes where the block size allows), fed to VEC successive
// MFMA steps.  The MFMA's k-slot q of lane group q then stands for k =
// kb + VEC q + v at step v: a permutation of the contraction index, applied
// to the LDS-held coefficient operand alike.
template <typename S, int B>
constexpr int rot_vec()
{
  constexpr int v = 16 / int(sizeof(S));
  return (B % v == 0) ? v : (B % 2 == 0 ? 2 : 1);
}
// k-steps whose operand loads are issued together, ahead of their products
constexpr int kRotUnroll = 8;

template <typename S, int VEC>
struct RotVec {
  using T = std::conditional_t<VEC == 1, S, S __attribute__((ext_vector_type(VEC)))>;
};

// Stores of one 16-row tile of the rotation: element e of lane (i16, q4) is
// row mfma_row(q4, e), column t*16 + i16 of the output.  Transposed across
// neighbouring lanes so that each lane writes 16 contiguous bytes of one row
// where the block size allows.
template <typename S, int B, int NT>
__device__ inline void rotate_store(const typename Mfma<S>::acc_t (&acc)[NT], size_t r0, int lane, int NC, S* __restrict__ Out, size_t ldv, size_t n)
{
  const int i16 = lane & 15, q4 = lane >> 4;
  using acc_t = typename Mfma<S>::acc_t;
  if constexpr (sizeof(S) == 4 && B % 4 == 0) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc_t r = acc[t];
      pair_transpose<kDppXor1, 1, 1>(r, 0, lane);
      pair_transpose<kDppXor1, 1, 1>(r, 2, lane);
      pair_transpose<kDppXor2, 2, 2>(r, 0, lane);
      pair_transpose<kDppXor2, 2, 2>(r, 1, lane);
      // r[j]: row 4 q4 + (i16 & 3), column t*16 + (i16 & ~3) + j
      const int jn = t * 16 + (i16 & ~3);
      const size_t row = r0 + 4 * q4 + (i16 & 3);
      if (jn < NC && row < n) {
        using V4 = S __attribute__((ext_vector_type(4)));
        *reinterpret_cast<V4*>(Out + size_t(jn / B) * ldv + row * B + (jn % B)) = V4{r[0], r[1], r[2], r[3]};
      }
    }
  }
  else if constexpr (sizeof(S) == 8 && B % 2 == 0) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc_t r = acc[t];
      pair_transpose<kDppXor1, 1, 1>(r, 0, lane);
      pair_transpose<kDppXor1, 1, 1>(r, 2, lane);
      // (r[e0], r[e0 + 1]), e0 = 0, 2: row mfma_row(q4, e0 + (i16 & 1)), columns t*16 + (i16 & ~1) + 0, 1
      const int jn = t * 16 + (i16 & ~1);
      if (jn < NC) {
        using V2 = S __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int e0 = 0; e0 < 4; e0 += 2) {
          const size_t row = r0 + mfma_row<S>(q4, e0 + (i16 & 1));
          if (row < n) *reinterpret_cast<V2*>(Out + size_t(jn / B) * ldv + row * B + (jn % B)) = V2{r[e0], r[e0 + 1]};
        }
      }
    }
  }
  else {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int jn = t * 16 + i16;
      if (jn >= NC) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const size_t row = r0 + mfma_row<S>(q4, e);
        if (row < n) Out[size_t(jn / B) * ldv + row * B + (jn % B)] = static_cast<S>(acc[t][e]);
      }
    }
  }
}

template <typename S, int B, int NT, int VEC>
__global__ __launch_bounds__(512) void k_rotate(const S* __restrict__ V, size_t ldv, int m, const double* __restrict__ Yg, int ybc, int jb0, int nb,
                                                S* __restrict__ Out, size_t n)
{
  using M = Mfma<S>;
  using acc_t = typename M::acc_t;
  constexpr int KS = 4 * VEC;                                     // k per super-step (VEC MFMA steps)
  constexpr int U = (kRotUnroll + VEC - 1) / VEC;                 // super-steps loaded together
  using VT = typename RotVec<S, VEC>::T;
  extern __shared__ unsigned char smem_raw[];
  S* Ys = reinterpret_cast<S*>(smem_raw);
  const int K = m * B, NC = nb * B;
  constexpr int ldy = NT * 16;
  for (int idx = threadIdx.x; idx < K * ldy; idx += blockDim.x) {
    const int kk = idx / ldy, jn = idx % ldy;
    double v = 0.0;
    if (jn < NC) {
      const int i = kk / B, ci = kk % B, j = jb0 + jn / B, cj = jn % B;
      v = Yg[((size_t(i) * ybc + j) * B + ci) * B + cj];
    }
    Ys[idx] = static_cast<S>(v);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const size_t ntiles = (n + 15) / 16;
  const int nsup = (K + KS - 1) / KS;
  for (size_t rt = size_t(blockIdx.x) * nwave + wave; rt < ntiles; rt += size_t(gridDim.x) * nwave) {
    const size_t r0 = rt * 16;
    const size_t ra = r0 + i16;
    const bool rok = ra < n;
    acc_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = acc_t{0, 0, 0, 0};
    for (int s0 = 0; s0 < nsup; s0 += U) {
      VT a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k0 = (s0 + u) * KS + VEC * q4; // VEC columns of one block row (VEC divides B)
        a[u] = VT{};
        if (rok && k0 < K) a[u] = *reinterpret_cast<const VT*>(V + size_t(k0 / B) * ldv + ra * B + (k0 % B));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (s0 + u >= nsup) break; // uniform over the wave
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const int kk = (s0 + u) * KS + VEC * q4 + v;
          const bool kok = kk < K;
          S av;
          if constexpr (VEC == 1)
            av = a[u];
          else
            av = a[u][v];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const S bv = kok ? Ys[kk * ldy + t * 16 + i16] : S(0);
            acc[t] = M::op(av, bv, acc[t]);
          }
        }
      }
    }
    rotate_store<S, B, NT>(acc, r0, lane, NC, Out, ldv, n);
  }
}

// rotate_store through a buffer resource over Out (element offsets from its
// base; rows past n and columns past NC move nothing)
template <typename S, int B, int NT>
__device__ inline void rotate_store_buf(const typename Mfma<S>::acc_t (&acc)[NT], size_t r0, int lane, int NC, __amdgpu_buffer_rsrc_t out, unsigned ldv,
                                        size_t n)
{
  const int i16 = lane & 15, q4 = lane >> 4;
  using acc_t = typename Mfma<S>::acc_t;
  if constexpr (sizeof(S) == 4 && B % 4 == 0) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc_t r = acc[t];
      pair_transpose<kDppXor1, 1, 1>(r, 0, lane);
      pair_transpose<kDppXor1, 1, 1>(r, 2, lane);
      pair_transpose<kDppXor2, 2, 2>(r, 0, lane);
      pair_transpose<kDppXor2, 2, 2>(r, 1, lane);
      const int jn = t * 16 + (i16 & ~3);
      const size_t row = r0 + 4 * q4 + (i16 & 3);
      const unsigned off = ((unsigned(jn / B) * ldv + unsigned(row) * B + unsigned(jn % B)) * 4u) | (unsigned((jn >= NC) | (row >= n)) << 31);
      buf_store(r, out, off);
    }
  }
  else if constexpr (sizeof(S) == 8 && B % 2 == 0) {
    using V2 = S __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc_t r = acc[t];
      pair_transpose<kDppXor1, 1, 1>(r, 0, lane);
      pair_transpose<kDppXor1, 1, 1>(r, 2, lane);
      const int jn = t * 16 + (i16 & ~1);
#pragma unroll
      for (int e0 = 0; e0 < 4; e0 += 2) {
        const size_t row = r0 + mfma_row<S>(q4, e0 + (i16 & 1));
        const unsigned off = ((unsigned(jn / B) * ldv + unsigned(row) * B + unsigned(jn % B)) * 8u) | (unsigned((jn >= NC) | (row >= n)) << 31);
        buf_store(V2{r[e0], r[e0 + 1]}, out, off);
      }
    }
  }
  else {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int jn = t * 16 + i16;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const size_t row = r0 + mfma_row<S>(q4, e);
        const unsigned off = ((unsigned(jn / B) * ldv + unsigned(row) * B + unsigned(jn % B)) * unsigned(sizeof(S))) | (unsigned((jn >= NC) | (row >= n)) << 31);
        buf_store(static_cast<S>(acc[t][e]), out, off);
      }
    }
  }
}

// The rotation for a short contraction (K = m*B of at most NSUP super-steps:
// the refinement sweeps, m = kc = 5): a tile's whole V row segment fits in
// NSUP vector registers per lane, so the wave loads the next tile's segment
// before this tile's products and stores.  Loads and stores go through
// range-checked buffer resources (V and Out each below 2^31 bytes), so no
// lane branches around them and a tile's stores are never waited for.  Y
// sits in LDS lane-major: the VEC*NT coefficients a lane multiplies in one
// super-step are contiguous (zeros past K), and with YREG all of them are
// held in registers for the whole launch.
template <typename S, int B, int NT, int VEC, int NSUP, bool YREG>
__device__ inline void rotate_short_body(const S* __restrict__ V, size_t ldv, int m, const double* __restrict__ Yg, int ybc, int jb0, int nb,
                                         S* __restrict__ Out, size_t n)
{
  using M = Mfma<S>;
  using acc_t = typename M::acc_t;
  constexpr int KS = 4 * VEC;
  constexpr int YL = VEC * NT; // coefficients per lane and super-step
  using VT = typename RotVec<S, VEC>::T;
  extern __shared__ unsigned char smem_raw[];
  S* Ys = reinterpret_cast<S*>(smem_raw);
  const int K = m * B, NC = nb * B;
  // Ys[((u*4 + q)*16 + i)*YL + v*NT + t] = Y(k = u*KS + VEC*q + v, column t*16 + i)
  for (int idx = threadIdx.x; idx < NSUP * 64 * YL; idx += blockDim.x) {
    const int g = idx / (16 * YL), rem = idx % (16 * YL);
    const int i = rem / YL, w = rem % YL, v = w / NT, t = w % NT;
    const int kk = (g >> 2) * KS + VEC * (g & 3) + v, jn = t * 16 + i;
    double y = 0.0;
    if (kk < K && jn < NC) {
      const int bi = kk / B, ci = kk % B, j = jb0 + jn / B, cj = jn % B;
      y = Yg[((size_t(bi) * ybc + j) * B + ci) * B + cj];
    }
    Ys[idx] = static_cast<S>(y);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const S* ylane = Ys + size_t(q4 * 16 + i16) * YL; // + u * 64 * YL
  S yr[YREG ? NSUP : 1][YL];
  if constexpr (YREG) {
#pragma unroll
    for (int u = 0; u < NSUP; ++u)
#pragma unroll
      for (int w = 0; w < YL; ++w) yr[u][w] = ylane[u * 64 * YL + w];
  }
  const size_t ntiles = (n + 15) / 16;
  const size_t stride = size_t(gridDim.x) * nwave;
  const __amdgpu_buffer_rsrc_t vr = buf_rsrc(V, size_t(m) * ldv * sizeof(S));
  const __amdgpu_buffer_rsrc_t orc = buf_rsrc(Out, size_t(nb) * ldv * sizeof(S));
  const unsigned ldv32 = unsigned(ldv);
  // the VEC columns of super-step u of tile t held by this lane (zero past n or K)
  // Offsets are formed without branches: the top bit, OR-ed in after the
  // additions, puts an access past the resource (every valid offset is
  // below 2^31).  Per super-step: the lane's column offset and its bit past K.
  unsigned koff[NSUP], kbad[NSUP];
#pragma unroll
  for (int u = 0; u < NSUP; ++u) {
    const int k0 = u * KS + VEC * q4;
    koff[u] = (unsigned(k0 / B) * ldv32 + unsigned(k0 % B)) * unsigned(sizeof(S));
    kbad[u] = unsigned(k0 >= K) << 31;
  }
  auto load_tile = [&](size_t t, VT (&a)[NSUP]) {
    const unsigned roff = (unsigned(t) * 16u + unsigned(i16)) * unsigned(B * sizeof(S));
    const unsigned rbad = unsigned(t * 16 + i16 >= n) << 31;
#pragma unroll
    for (int u = 0; u < NSUP; ++u) a[u] = buf_load<VT>(vr, (koff[u] + roff) | kbad[u] | rbad);
  };
  // one tile: products of its row segment a, then its stores
  auto tile = [&](size_t t, const VT (&a)[NSUP]) {
    acc_t acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < NSUP; ++u) {
      S yu[YL];
#pragma unroll
      for (int w = 0; w < YL; ++w) {
        if constexpr (YREG)
          yu[w] = yr[u][w];
        else
          yu[w] = ylane[u * 64 * YL + w];
      }
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        S av;
        if constexpr (VEC == 1)
          av = a[u];
        else
          av = a[u][v];
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = M::op(av, yu[v * NT + j], acc[j]);
      }
    }
    rotate_store_buf<S, B, NT>(acc, t * 16, lane, NC, orc, ldv32, n);
  };
  // Two register sets in turn: the loads of the next tile are issued (and
  // kept ahead of this tile's products by a scheduling barrier) before the
  // products and stores of the current one.  Both halves run unconditionally:
  // a tile past the last one reads zeros and stores nothing.
  size_t rt = size_t(blockIdx.x) * nwave + size_t(__builtin_amdgcn_readfirstlane(wave));
  VT ra_[NSUP], rb_[NSUP];
  load_tile(rt, ra_);
  for (; rt < ntiles; rt += 2 * stride) {
    load_tile(rt + stride, rb_);
    __builtin_amdgcn_sched_barrier(0);
    tile(rt, ra_);
    load_tile(rt + 2 * stride, ra_);
    __builtin_amdgcn_sched_barrier(0);
    tile(rt + stride, rb_);
  }
}
template <typename S, int B, int NT, int VEC, int NSUP, bool YREG>
__global__ __launch_bounds__(512) void k_rotate_short(const S* __restrict__ V, size_t ldv, int m, const double* __restrict__ Yg, int ybc, int jb0, int nb,
                                                      S* __restrict__ Out, size_t n)
{
  rotate_short_body<S, B, NT, VEC, NSUP, YREG>(V, ldv, m, Yg, ybc, jb0, nb, Out, n);
}
template <typename S, int B, int NT, int VEC, int NSUP, bool YREG>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5))) void k_rotate_short_w5(const S* __restrict__ V, size_t ldv, int m, const double* __restrict__ Yg, int ybc, int jb0, int nb,
                                                      S* __restrict__ Out, size_t n)
{
  rotate_short_body<S, B, NT, VEC, NSUP, YREG>(V, ldv, m, Yg, ybc, jb0, nb, Out, n);
}

// k-steps per prefetch group of the long-contraction rotation
constexpr int kRotGroupK = 32;

// The rotation for a long contraction (the thick restart, K = m*B in the
// hundreds): a tile's operand row streams through two register sets in
// groups of kRotGroupK k-steps, the loads of the next group (of this tile,
// or the first of the next tile) issued before the products of the current
// one.  Loads and stores are range-checked buffer accesses (zeros past n and
// K, nothing stored past n or NC); a tile's stores are issued during the
// next tile's first group, so no product waits for a store.  Y sits in LDS
// lane-major over ng2 groups (an even count, zeros past K).
template <typename S, int B, int NT, int VEC>
__global__ __launch_bounds__(512) void k_rotate_long(const S* __restrict__ V, size_t ldv, int m, const double* __restrict__ Yg, int ybc, int jb0, int nb,
                                                     S* __restrict__ Out, size_t n, int ng2)
{
  using M = Mfma<S>;
  using acc_t = typename M::acc_t;
  constexpr int KS = 4 * VEC;
  constexpr int U = kRotGroupK / KS; // super-steps per group
  constexpr int YL = VEC * NT;
  using VT = typename RotVec<S, VEC>::T;
  extern __shared__ unsigned char smem_raw[];
  S* Ys = reinterpret_cast<S*>(smem_raw);
  const int K = m * B, NC = nb * B;
  // Ys[((u*4 + q)*16 + i)*YL + v*NT + t] = Y(k = u*KS + VEC*q + v, column t*16 + i), u over all super-steps
  for (int idx = threadIdx.x; idx < ng2 * U * 64 * YL; idx += blockDim.x) {
    const int g = idx / (16 * YL), rem = idx % (16 * YL);
    const int i = rem / YL, w = rem % YL, v = w / NT, t = w % NT;
    const int kk = (g >> 2) * KS + VEC * (g & 3) + v, jn = t * 16 + i;
    double y = 0.0;
    if (kk < K && jn < NC) {
      const int bi = kk / B, ci = kk % B, j = jb0 + jn / B, cj = jn % B;
      y = Yg[((size_t(bi) * ybc + j) * B + ci) * B + cj];
    }
    Ys[idx] = static_cast<S>(y);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const S* ylane = Ys + size_t(q4 * 16 + i16) * YL;
  const size_t ntiles = (n + 15) / 16;
  const size_t stride = size_t(gridDim.x) * nwave;
  const __amdgpu_buffer_rsrc_t orc = buf_rsrc(Out, size_t(nb) * ldv * sizeof(S));
  const unsigned ldv32 = unsigned(ldv);
  // A super-step's columns lie in at most kSpan consecutive blocks of V; each
  // super-step reads through a resource over those blocks alone (below 2^31
  // bytes however many blocks V has).  Offsets are formed without branches:
  // the top bit, OR-ed in after the additions, puts an access past the
  // resource (every valid offset is below 2^31).
  constexpr int kSpan = (KS + B - 1) / B + 1;
  auto load_group = [&](size_t t, int g, VT (&a)[U]) {
    const unsigned roff = (unsigned(t) * 16u + unsigned(i16)) * unsigned(B * sizeof(S));
    const unsigned rbad = unsigned(t * 16 + i16 >= n) << 31;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s0 = (g * U + u) * KS, blk0 = s0 / B;
      const int span = (m - blk0) < kSpan ? (m - blk0) : kSpan;
      const __amdgpu_buffer_rsrc_t vr = buf_rsrc_uniform(V + size_t(blk0) * ldv, size_t(span > 0 ? span : 0) * ldv * sizeof(S));
      const int k0 = s0 + VEC * q4;
      const unsigned koff = (unsigned(k0 / B - blk0) * ldv32 + unsigned(k0 % B)) * unsigned(sizeof(S));
      a[u] = buf_load<VT>(vr, (koff + roff) | (unsigned(k0 >= K) << 31) | rbad);
    }
  };
  auto mma_group = [&](int g, const VT (&a)[U], acc_t (&acc)[NT]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      S yu[YL];
#pragma unroll
      for (int w = 0; w < YL; ++w) yu[w] = ylane[(g * U + u) * 64 * YL + w];
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        S av;
        if constexpr (VEC == 1)
          av = a[u];
        else
          av = a[u][v];
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = M::op(av, yu[v * NT + j], acc[j]);
      }
    }
  };
  // the next group after (t, g): (t, g + 1), or the first of the next tile
  auto next_t = [&](size_t t, int g) { return g + 1 >= ng2 ? t + stride : t; };
  auto next_g = [&](int g) { return g + 1 >= ng2 ? 0 : g + 1; };
  size_t rt = size_t(blockIdx.x) * nwave + size_t(__builtin_amdgcn_readfirstlane(wave));
  VT A[U], Bq[U];
  acc_t prev[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) prev[j] = acc_t{0, 0, 0, 0};
  size_t prev_t = ntiles; // rows past n: the first tile's "previous" stores move nothing
  load_group(rt, 0, A);
  for (; rt < ntiles; rt += stride) {
    acc_t acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = acc_t{0, 0, 0, 0};
    // first pair of groups, the previous tile's stores after the prefetch
    load_group(rt, 1, Bq);
    rotate_store_buf<S, B, NT>(prev, prev_t * 16, lane, NC, orc, ldv32, n);
    __builtin_amdgcn_sched_barrier(0);
    mma_group(0, A, acc);
    load_group(next_t(rt, 1), next_g(1), A);
    __builtin_amdgcn_sched_barrier(0);
    mma_group(1, Bq, acc);
    for (int g = 2; g < ng2; g += 2) {
      load_group(rt, g + 1, Bq);
      __builtin_amdgcn_sched_barrier(0);
      mma_group(g, A, acc);
      load_group(next_t(rt, g + 1), next_g(g + 1), A);
      __builtin_amdgcn_sched_barrier(0);
      mma_group(g + 1, Bq, acc);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) prev[j] = acc[j];
    prev_t = rt;
  }
  rotate_store_buf<S, B, NT>(prev, prev_t * 16, lane, NC, orc, ldv32, n);
}

inline int pick_tpw(int per_wave)
{
  if (per_wave <= 6) return per_wave < 1 ? 1 : per_wave;
  if (per_wave <= 8) return 8;
  if (per_wave <= 12) return 12;
  return 16;
}

// Queue the Gram partials for blocks [0, k) of V against W; returns the
// number of partial rows written (nwg * RS) in *nw.
// partials: where the rows go; nullptr = the context's growable buffer
// (ctx->partials, enlarged here when needed), else a fixed buffer of cap
// doubles, for which the workgroup count is capped so that the rows fit.
// tiles per wave up to which the Gram takes its buffer form (two register
// sets of kU * (TPW + 2) values per lane)
constexpr int kGramBufTpw = 4;

template <typename S, int B>
int queue_gram_partials(trl_hip_ctx* ctx, size_t n, int k, const S* V, const S* W, const S* md, int* nw, const int* gate, double* partials = nullptr,
                        size_t cap = 0)
{
  constexpr int BP = pad_b(B), BPT = 16 / BP;
  const int ntile = (k + BPT - 1) / BPT;
  const int TG = ntile >= 4 ? 4 : (ntile >= 2 ? 2 : 1);
  const int RS = 4 / TG;
  const int tpw = pick_tpw((ntile + TG - 1) / TG);
  if ((ntile + TG - 1) / TG > 16) return TRL_HIP_EINVAL;
  const size_t ngroups = ((n + 3) / 4 + kU - 1) / kU;
  size_t g = (ngroups + RS - 1) / RS;
  const size_t gmax = size_t(kGramWgPerCu) * ctx->num_cus;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  const size_t P = size_t(k) * B * B;
  if (partials) {
    const size_t fit = cap / (size_t(RS) * P);
    if (fit < 1) return TRL_HIP_EINVAL;
    if (g > fit) g = fit;
  }
  else {
    int rc = trl_hip_impl::ensure_partials(ctx, g * RS * P);
    if (rc) return rc;
    partials = ctx->partials;
  }
  const size_t ldv = n * B;
  const dim3 grid{unsigned(g)}, blk{unsigned(kThreads)};
  // buffer form (FP32: 16 blocks of the 160^3 stencil, b = 8, 485 -> 440 us;
  // FP64 over 32 blocks 1.56 ms either way, kept on k_gram): W (and md) and
  // each tile's BPT blocks of V below 2^31 bytes
  const size_t lim = size_t(1) << 31;
  const bool buf = sizeof(S) == 4 && tpw <= kGramBufTpw && size_t(BPT) * ldv * sizeof(S) < lim && ldv * sizeof(S) < lim;
#define TRL_GRAM(TPW_)                                                                                                                             \
  do {                                                                                                                                             \
    if (buf)                                                                                                                                       \
      hipLaunchKernelGGL((k_gram_buf<S, B, TPW_>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, partials, gate);                            \
    else                                                                                                                                           \
      hipLaunchKernelGGL((k_gram<S, B, TPW_>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, partials, gate);                                \
  } while (0)
  switch (tpw) {
  case 1:
    TRL_GRAM(1);
    break;
  case 2:
    TRL_GRAM(2);
    break;
  case 3:
    TRL_GRAM(3);
    break;
  case 4:
    TRL_GRAM(4);
    break;
  case 5:
    TRL_GRAM(5);
    break;
  case 6:
    TRL_GRAM(6);
    break;
  case 8:
    TRL_GRAM(8);
    break;
  case 12:
    TRL_GRAM(12);
    break;
  default:
    TRL_GRAM(16);
    break;
  }
#undef TRL_GRAM
  TRL_HIP_LAUNCH_CHECK();
  *nw = int(g * RS);
  return 0;
}

// largest block count per Gram launch: tiles per wave <= 16 and k*b*b <= 16384
template <int B>
constexpr int gram_max_k()
{
  constexpr int BP = pad_b(B), BPT = 16 / BP;
  constexpr int by_tiles = 4 * 16 * BPT;
  constexpr int by_p = 16384 / (B * B);
  return by_tiles < by_p ? by_tiles : by_p;
}

} // namespace

namespace trl_hip_impl {

// Second-pass test of classical Gram-Schmidt with reorthogonalisation
// (Daniel, Gragg, Kaufman and Stewart, Math. Comp. 30 (1976)): C holds the
// first pass's coefficients V_q^T M W (blocks q < k) and W^T M W (block k).
// Column j needs a second pass when its norm after the first pass,
// ||w_j||^2 - sum_q ||c_qj||^2, falls below eta2 ||w_j||^2.  gate = 1 if any
// column does.
template <int B>
__global__ void k_dgks(const double* __restrict__ C, int k, double eta2, int* __restrict__ gate)
{
  __shared__ int need;
  if (threadIdx.x == 0) need = 0;
  __syncthreads();
  const int j = threadIdx.x;
  if (j < B) {
    const double w2 = C[size_t(k) * B * B + j * B + j];
    double c2 = 0.0;
    for (int q = 0; q < k; ++q)
      for (int c = 0; c < B; ++c) {
        const double x = C[size_t(q) * B * B + c * B + j];
        c2 = fma(x, x, c2);
      }
    if (!(w2 - c2 >= eta2 * w2) || !(w2 > 0.0)) atomicOr(&need, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *gate = need;
}

int launch_dgks(trl_hip_ctx* ctx, int b, int k, const double* C, double eta2, int* gate)
{
  switch (b) {
  case 1:
    hipLaunchKernelGGL(k_dgks<1>, dim3(1), dim3(64), 0, ctx->stream, C, k, eta2, gate);
    break;
  case 2:
    hipLaunchKernelGGL(k_dgks<2>, dim3(1), dim3(64), 0, ctx->stream, C, k, eta2, gate);
    break;
  case 4:
    hipLaunchKernelGGL(k_dgks<4>, dim3(1), dim3(64), 0, ctx->stream, C, k, eta2, gate);
    break;
  case 6:
    hipLaunchKernelGGL(k_dgks<6>, dim3(1), dim3(64), 0, ctx->stream, C, k, eta2, gate);
    break;
  case 8:
    hipLaunchKernelGGL(k_dgks<8>, dim3(1), dim3(64), 0, ctx->stream, C, k, eta2, gate);
    break;
  case 16:
    hipLaunchKernelGGL(k_dgks<16>, dim3(1), dim3(64), 0, ctx->stream, C, k, eta2, gate);
    break;
  default:
    return TRL_HIP_EINVAL;
  }
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_gram(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const void* W, const void* md, double* C, const int* gate,
                double* partials, size_t cap)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const S* v = static_cast<const S*>(V);
      const S* w = static_cast<const S*>(W);
      const S* m = static_cast<const S*>(md);
      if (n == 0) {
        hipLaunchKernelGGL(k_reduce, dim3(unsigned(size_t(k) * B * B)), dim3(kThreads), 0, ctx->stream, partials ? partials : ctx->partials, 0, k * B * B,
                           C, gate);
        TRL_HIP_LAUNCH_CHECK();
        return 0;
      }
      constexpr int kmax = gram_max_k<B>();
      for (int j0 = 0; j0 < k; j0 += kmax) {
        const int kc = (k - j0) < kmax ? (k - j0) : kmax;
        int nw = 0;
        int rc = queue_gram_partials<S, B>(ctx, n, kc, v + size_t(j0) * n * B, w, m, &nw, gate, partials, cap);
        if (rc) return rc;
        const int P = kc * B * B;
        hipLaunchKernelGGL(k_reduce, dim3(unsigned(P)), dim3(kThreads), 0, ctx->stream, partials ? partials : ctx->partials, nw, P, C + size_t(j0) * B * B,
                           gate);
        TRL_HIP_LAUNCH_CHECK();
      }
    });
  });
  return 0;
}

// ---- the allocation-free Gram entry point's scratch --------------------------
//
// A file-scope device array holds the per-workgroup partial sums of
// trl_hip_gram_noalloc.  Each device holds its own copy from the moment the
// code object is loaded until the process ends.  In host code `scratch` is
// only a stand-in for the symbol: its device address is resolved once per
// device (hipGetSymbolAddress, with that device current) and that address is
// what the kernels receive.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
} // namespace

int noalloc_scratch(int device, void** out)
{
  if (!out || device < 0 || device >= kMaxDevices) return TRL_HIP_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    TRL_HIP_TRY(hipGetDevice(&cur));
    if (cur != device) return TRL_HIP_EINVAL; // the device must be current
    void* ptr = nullptr;
    TRL_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return TRL_HIP_ENOMEM;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

int launch_update(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const double* C, void* W, const int* gate)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const int kmax = int((48 * 1024) / (B * B * sizeof(S)));
      size_t g = (n + kThreads - 1) / kThreads;
      const size_t gmax = size_t(8) * ctx->num_cus;
      if (g > gmax) g = gmax;
      for (int j0 = 0; j0 < k; j0 += kmax) {
        const int kc = (k - j0) < kmax ? (k - j0) : kmax;
        const size_t shm = size_t(kc) * B * B * sizeof(S);
        hipLaunchKernelGGL((k_update<S, B>), dim3(unsigned(g)), dim3(kThreads), shm, ctx->stream, static_cast<const S*>(V) + size_t(j0) * n * B, n * B, kc,
                           C + size_t(j0) * B * B, static_cast<S*>(W), n, gate);
        TRL_HIP_LAUNCH_CHECK();
      }
    });
  });
  return 0;
}

// Cholesky of a reduced Gram matrix G (B x B, global) by one thread.
template <int B>
__global__ void k_chol_small(const double* __restrict__ G, const double* __restrict__ Rprev, double* __restrict__ Rout, double* __restrict__ Rinv,
                             int* __restrict__ flags)
{
  __shared__ double Gs[B * B];
  __shared__ double Rl[16][16], Xl[16][16];
  for (int i = threadIdx.x; i < B * B; i += blockDim.x) Gs[i] = G[i];
  __syncthreads();
  if constexpr (B <= 8)
    chol_lanes<B>(Gs, Rprev, Rout, Rinv, flags, &Rl[0][0], &Xl[0][0]);
  else if (threadIdx.x == 0)
    chol_bxb(Gs, B, Rprev, Rout, Rinv, flags, Rl, Xl);
}

// Rows of one Gram pass of CholQR2: G = Y^T M Y in FP64 products, Y = W
// (TRANSFORM false) or Y = W R1^{-1} rounded to the panel type and formed
// on the fly from W (TRANSFORM: the second pass, Rinv = R1^{-1}).  Each
// thread loads four rows (grid-stride) before their arithmetic, so four
// row loads per lane are in flight; it keeps the B(B+1)/2 entries of the
// upper triangle, and the workgroup sums them (butterfly per wave, waves in
// order) into its packed partial Gram, entry-major (entry i of workgroup w
// at i*G + w) for k_chol_partials.
template <typename S, int B, bool TRANSFORM>
__device__ inline void gram_row_add(const S (&w)[B], double m, const double* Rs, double (&g)[B * (B + 1) / 2])
{
  double y[B];
  if constexpr (TRANSFORM) {
    double acc[B];
#pragma unroll
    for (int c = 0; c < B; ++c) acc[c] = 0.0;
#pragma unroll
    for (int cp = 0; cp < B; ++cp) {
      const double x = static_cast<double>(w[cp]);
#pragma unroll
      for (int c = cp; c < B; ++c) acc[c] = fma(x, Rs[cp * B + c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < B; ++c) y[c] = static_cast<double>(static_cast<S>(acc[c]));
  }
  else {
#pragma unroll
    for (int c = 0; c < B; ++c) y[c] = static_cast<double>(w[c]);
  }
  int idx = 0;
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const double yi = y[i] * m;
#pragma unroll
    for (int jj = i; jj < B; ++jj) {
      g[idx] = fma(yi, y[jj], g[idx]);
      ++idx;
    }
  }
}

template <typename S, int B, bool TRANSFORM>
__global__ __launch_bounds__(kThreads) void k_gram_rows(const S* __restrict__ W, const S* __restrict__ md, size_t n, const double* __restrict__ Rinv,
                                                        double* __restrict__ partials)
{
  constexpr int NP = B * (B + 1) / 2;
  constexpr int U = 4;
  __shared__ double Rs[TRANSFORM ? B * B : 1];
  __shared__ double red[kThreads / 64][NP];
  if constexpr (TRANSFORM) {
    for (int i = threadIdx.x; i < B * B; i += blockDim.x) Rs[i] = Rinv[i];
    __syncthreads();
  }
  double g[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) g[i] = 0.0;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  size_t row = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; row + (U - 1) * stride < n; row += U * stride) {
    S w[U][B];
#pragma unroll
    for (int u = 0; u < U; ++u) RowIO<S, B>::template load<false>(W + (row + u * stride) * B, w[u]);
    double mm[U];
#pragma unroll
    for (int u = 0; u < U; ++u) mm[u] = md ? static_cast<double>(ld<false>(md + row + u * stride)) : 1.0;
#pragma unroll
    for (int u = 0; u < U; ++u) gram_row_add<S, B, TRANSFORM>(w[u], mm[u], Rs, g);
  }
  for (; row < n; row += stride) {
    S w[B];
    RowIO<S, B>::template load<false>(W + row * B, w);
    gram_row_add<S, B, TRANSFORM>(w, md ? static_cast<double>(ld<false>(md + row)) : 1.0, Rs, g);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    double v = g[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  const unsigned G = gridDim.x;
  for (int i = threadIdx.x; i < NP; i += blockDim.x) {
    double t = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) t += red[q][i];
    partials[size_t(i) * G + blockIdx.x] = t;
  }
}

// The end of a Gram-by-rows pass: one workgroup of kReduceThreads sums the
// G workgroups' packed partials -- wave q takes the entries i = q, q + 16,
// ..., each lane the workgroups w = lane, lane + 64, ... of an entry, then a
// butterfly (a fixed order) --, mirrors the sums into the b x b Gram and
// factors it (chol_lanes): Rout = R Rprev (or R), Rinv = R^{-1}, a
// breakdown into flags.
constexpr int kReduceThreads = 1024;

template <int B>
__global__ __launch_bounds__(kReduceThreads) void k_chol_partials(const double* __restrict__ partials, int G, const double* __restrict__ Rprev,
                                                                  double* __restrict__ Rout, double* __restrict__ Rinv, int* __restrict__ flags)
{
  static_assert(B <= 8, "chol_lanes");
  constexpr int NP = B * (B + 1) / 2;
  constexpr int NW = kReduceThreads / 64;
  __shared__ double Gs[B * B], Rsh[B * B], T[B];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = wave; i < NP; i += NW) {
    const double* src = partials + size_t(i) * G;
    double a0 = 0.0, a1 = 0.0;
    int w = lane;
    for (; w + 64 < G; w += 128) {
      a0 += src[w];
      a1 += src[w + 64];
    }
    if (w < G) a0 += src[w];
    double x = a0 + a1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) {
      int r = 0, rem = i;
      while (rem >= B - r) {
        rem -= B - r;
        ++r;
      }
      const int c = r + rem;
      Gs[r * B + c] = x;
      Gs[c * B + r] = x;
    }
  }
  __syncthreads();
  chol_lanes<B>(Gs, Rprev, Rout, Rinv, flags, Rsh, T);
}

// W <- (W R1^{-1}) R2^{-1} with the intermediate rounded to the panel type:
// the two scale passes of CholQR2 in one read and one write of W.
template <typename S, int B>
__global__ __launch_bounds__(kThreads) void k_scale2(S* __restrict__ W, const double* __restrict__ R1inv, const double* __restrict__ R2inv, size_t n)
{
  __shared__ double Rs[2][B * B];
  for (int i = threadIdx.x; i < B * B; i += blockDim.x) {
    Rs[0][i] = R1inv[i];
    Rs[1][i] = R2inv[i];
  }
  __syncthreads();
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t row = size_t(blockIdx.x) * blockDim.x + threadIdx.x; row < n; row += stride) {
    S w[B];
    RowIO<S, B>::template load<true>(W + row * B, w);
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      double acc[B];
#pragma unroll
      for (int c = 0; c < B; ++c) acc[c] = 0.0;
#pragma unroll
      for (int cp = 0; cp < B; ++cp) {
        const double x = static_cast<double>(w[cp]);
#pragma unroll
        for (int c = cp; c < B; ++c) acc[c] = fma(x, Rs[pass][cp * B + c], acc[c]);
      }
#pragma unroll
      for (int c = 0; c < B; ++c) w[c] = static_cast<S>(acc[c]);
    }
    RowIO<S, B>::template store<true>(W + row * B, w);
  }
}

// Cholesky-QR for b <= 8 by rows.  CholQR (passes = 1) in two panel
// passes: G = W^T M W (read W), R; W <- W R^{-1} (read and write W).
// CholQR2 in four: G1 = W^T M W (read W), R1; G2 = Y^T M Y for Y = W R1^{-1}
// formed on the fly (read W), R2 and R = R2 R1; W <- Y R2^{-1} (read and
// write W) -- where two CholQR passes read W three times and write it
// twice.  Per Gram a thread-per-row FP64 kernel and one workgroup that sums
// its partials and factors the b x b Gram (k_chol_partials).
int launch_cholqr_rows(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, const void* md, double* R, int passes)
{
  double* R1 = ctx->small;
  double* Rinv1 = ctx->small + 256;
  double* Rinv2 = ctx->small + 512;
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      if constexpr (B <= 8) {
        S* w = static_cast<S*>(W);
        const S* m = static_cast<const S*>(md);
        size_t g = (n + kThreads - 1) / kThreads;
        const size_t gmax = size_t(2) * ctx->num_cus;
        if (g > gmax) g = gmax;
        if (g < 1) g = 1;
        constexpr int NP = B * (B + 1) / 2;
        int rc = ensure_partials(ctx, g * NP);
        if (rc) return rc;
        const dim3 grid{unsigned(g)}, blk{unsigned(kThreads)};
        hipLaunchKernelGGL((k_gram_rows<S, B, false>), grid, blk, 0, ctx->stream, w, m, n, nullptr, ctx->partials);
        TRL_HIP_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_chol_partials<B>), dim3(1), dim3(kReduceThreads), 0, ctx->stream, ctx->partials, int(g), nullptr, passes == 1 ? R : R1,
                           Rinv1, ctx->flags);
        TRL_HIP_LAUNCH_CHECK();
        if (passes == 2) {
          hipLaunchKernelGGL((k_gram_rows<S, B, true>), grid, blk, 0, ctx->stream, w, m, n, Rinv1, ctx->partials);
          TRL_HIP_LAUNCH_CHECK();
          hipLaunchKernelGGL((k_chol_partials<B>), dim3(1), dim3(kReduceThreads), 0, ctx->stream, ctx->partials, int(g), R1, R, Rinv2, ctx->flags);
          TRL_HIP_LAUNCH_CHECK();
        }
        if (n == 0) return 0;
        size_t gs = (n + kThreads - 1) / kThreads;
        const size_t gsmax = size_t(8) * ctx->num_cus;
        if (gs > gsmax) gs = gsmax;
        if (passes == 2 && sizeof(S) == 8) {
          // FP64: W R1^-1 R2^-1 as one product with the FP64 triangle R1^-1 R2^-1
          // (no intermediate rounding to skip): 88 us against 163 us for the
          // two-triangle pass, whose 72 factors held in registers left two
          // waves per SIMD (one FP64 panel of the 160^3 stencil, b = 8)
          double* Rinv12 = ctx->small + 768;
          const int rc2 = launch_small(ctx, 3, B, Rinv1, Rinv2, Rinv12, nullptr);
          if (rc2) return rc2;
          hipLaunchKernelGGL((k_scale<S, B>), dim3(unsigned(gs)), dim3(kThreads), 0, ctx->stream, w, Rinv12, n);
        }
        else if (passes == 2)
          hipLaunchKernelGGL((k_scale2<S, B>), dim3(unsigned(gs)), dim3(kThreads), 0, ctx->stream, w, Rinv1, Rinv2, n);
        else
          hipLaunchKernelGGL((k_scale<S, B>), dim3(unsigned(gs)), dim3(kThreads), 0, ctx->stream, w, Rinv1, n);
        TRL_HIP_LAUNCH_CHECK();
      }
      else
        return TRL_HIP_EINVAL;
    });
  });
  return 0;
}

// Cholesky-QR with the Gram matrix on MFMA (the block-product kernel on W
// against itself), a one-thread FP64 Cholesky and the FP64 scale; one panel
// read for the Gram and one read + write for the scale per pass.  Blocks of
// b <= 8 take the rows form above.
int launch_cholqr_gram(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, const void* md, double* R, int passes)
{
  if (b <= 8 && (passes == 1 || passes == 2)) return launch_cholqr_rows(ctx, t, n, b, W, md, R, passes);
  double* R1 = ctx->small;
  double* Rinv = ctx->small + 256;
  double* G = ctx->small + 768;
  for (int pass = 0; pass < passes; ++pass) {
    int rc = launch_gram(ctx, t, n, b, 1, W, W, md, G);
    if (rc) return rc;
    const bool last = pass == passes - 1;
    if (n == 0 && pass > 0) continue;
    TRL_DISPATCH_T(t, {
      TRL_DISPATCH_B(b, {
        hipLaunchKernelGGL((k_chol_small<B>), dim3(1), dim3(64), 0, ctx->stream, G, pass == 0 ? nullptr : R1, last ? R : R1, Rinv, ctx->flags);
        TRL_HIP_LAUNCH_CHECK();
        if (n == 0) return 0;
        size_t gs = (n + kThreads - 1) / kThreads;
        const size_t gsmax = size_t(8) * ctx->num_cus;
        if (gs > gsmax) gs = gsmax;
        hipLaunchKernelGGL((k_scale<S, B>), dim3(unsigned(gs)), dim3(kThreads), 0, ctx->stream, static_cast<S*>(W), Rinv, n);
        TRL_HIP_LAUNCH_CHECK();
      });
    });
  }
  return 0;
}

// Cholesky-QR in the inner product of a sparse SPD B, given BW = B W: per
// pass G = W^T (BW) on MFMA, the FP64 Cholesky, and W, BW <- W R^{-1},
// BW R^{-1} (so BW stays B times the new W without another product).
int launch_cholqr_bprod(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, void* BW, double* R, int passes)
{
  double* R1 = ctx->small;
  double* Rinv = ctx->small + 256;
  double* G = ctx->small + 768;
  for (int pass = 0; pass < passes; ++pass) {
    int rc = launch_gram(ctx, t, n, b, 1, W, BW, nullptr, G);
    if (rc) return rc;
    const bool last = pass == passes - 1;
    TRL_DISPATCH_T(t, {
      TRL_DISPATCH_B(b, {
        hipLaunchKernelGGL((k_chol_small<B>), dim3(1), dim3(64), 0, ctx->stream, G, pass == 0 ? nullptr : R1, last ? R : R1, Rinv, ctx->flags);
        TRL_HIP_LAUNCH_CHECK();
        if (n == 0) return 0;
        size_t gs = (n + kThreads - 1) / kThreads;
        const size_t gsmax = size_t(8) * ctx->num_cus;
        if (gs > gsmax) gs = gsmax;
        hipLaunchKernelGGL((k_scale<S, B>), dim3(unsigned(gs)), dim3(kThreads), 0, ctx->stream, static_cast<S*>(W), Rinv, n);
        TRL_HIP_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_scale<S, B>), dim3(unsigned(gs)), dim3(kThreads), 0, ctx->stream, static_cast<S*>(BW), Rinv, n);
        TRL_HIP_LAUNCH_CHECK();
      });
    });
  }
  return 0;
}

int launch_cholqr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, const void* md, double* R, int passes)
{
  // default: the Gram on MFMA (launch_cholqr_gram), measured faster than the
  // fused thread-per-row passes below; TRL_HIP_CHOLQR=fused selects those
  static const bool fused = [] {
    const char* e = std::getenv("TRL_HIP_CHOLQR");
    return e && std::string(e) == "fused";
  }();
  if (!fused) return launch_cholqr_gram(ctx, t, n, b, W, md, R, passes);
  // scratch: [0,256) R1, [256,512) Rinv of pass 1, [512,768) Rinv of pass 2
  double* R1 = ctx->small;
  double* Rinv1 = ctx->small + 256;
  double* Rinv2 = ctx->small + 512;
  unsigned* counter = reinterpret_cast<unsigned*>(ctx->flags + 8);
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      S* w = static_cast<S*>(W);
      const S* m = static_cast<const S*>(md);
      constexpr int NP = B * (B + 1) / 2;
      size_t g = (n + kThreads - 1) / kThreads;
      const size_t gmax = size_t(kGramWgPerCu) * ctx->num_cus;
      if (g > gmax) g = gmax;
      if (g < 1) g = 1;
      int rc = ensure_partials(ctx, g * NP);
      if (rc) return rc;
      const dim3 grid{unsigned(g)}, blk{unsigned(kThreads)};
      hipLaunchKernelGGL((k_cholqr_pass<S, B, false>), grid, blk, 0, ctx->stream, w, m, n, nullptr, ctx->partials, counter, nullptr, passes == 1 ? R : R1,
                         Rinv1, ctx->flags);
      TRL_HIP_LAUNCH_CHECK();
      const double* last_inv = Rinv1;
      if (passes == 2) {
        hipLaunchKernelGGL((k_cholqr_pass<S, B, true>), grid, blk, 0, ctx->stream, w, m, n, Rinv1, ctx->partials, counter, R1, R, Rinv2, ctx->flags);
        TRL_HIP_LAUNCH_CHECK();
        last_inv = Rinv2;
      }
      if (n > 0) {
        size_t gs = (n + kThreads - 1) / kThreads;
        const size_t gsmax = size_t(8) * ctx->num_cus;
        if (gs > gsmax) gs = gsmax;
        hipLaunchKernelGGL((k_scale<S, B>), dim3(unsigned(gs)), dim3(kThreads), 0, ctx->stream, w, last_inv, n);
        TRL_HIP_LAUNCH_CHECK();
      }
    });
  });
  return 0;
}

// super-steps of the short-contraction rotation (register cost 2 NSUP vectors per lane)
constexpr int kRotShortSteps = 6;

// coefficient registers per lane up to which the short rotation holds Y in registers
constexpr int kRotYRegBytes = 192;

template <typename S, int B, int VS, int NT, int NSUP>
void launch_rotate_short_nt(trl_hip_ctx* ctx, size_t g, const S* v, size_t n, int m, const double* Y, int ybc, int j0, int nb, S* o)
{
  constexpr bool yreg = NSUP * VS * NT * int(sizeof(S)) <= kRotYRegBytes;
  const size_t shm = size_t(NSUP) * 64 * VS * NT * sizeof(S);
  // FP32 with Y in registers: the five-wave register budget (FP32 sweep
  // rotation, m = kc = 5 on the 160^3 stencil: 266 us against 272 us)
  if constexpr (sizeof(S) == 4 && yreg)
    hipLaunchKernelGGL((k_rotate_short_w5<S, B, NT, VS, NSUP, yreg>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
  else
    hipLaunchKernelGGL((k_rotate_short<S, B, NT, VS, NSUP, yreg>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
}

template <typename S, int B, int VS>
int launch_rotate_short(trl_hip_ctx* ctx, int nt, int nsup, size_t g, const S* v, size_t n, int m, const double* Y, int ybc, int j0, int nb,
                        S* o)
{
  auto by_steps = [&](auto NTc) {
    constexpr int NT = decltype(NTc)::value;
    switch (nsup) {
    case 1:
      launch_rotate_short_nt<S, B, VS, NT, 1>(ctx, g, v, n, m, Y, ybc, j0, nb, o);
      break;
    case 2:
      launch_rotate_short_nt<S, B, VS, NT, 2>(ctx, g, v, n, m, Y, ybc, j0, nb, o);
      break;
    case 3:
      launch_rotate_short_nt<S, B, VS, NT, 3>(ctx, g, v, n, m, Y, ybc, j0, nb, o);
      break;
    case 4:
      launch_rotate_short_nt<S, B, VS, NT, 4>(ctx, g, v, n, m, Y, ybc, j0, nb, o);
      break;
    case 5:
      launch_rotate_short_nt<S, B, VS, NT, 5>(ctx, g, v, n, m, Y, ybc, j0, nb, o);
      break;
    default:
      launch_rotate_short_nt<S, B, VS, NT, 6>(ctx, g, v, n, m, Y, ybc, j0, nb, o);
      break;
    }
  };
  switch (nt) {
  case 1:
    by_steps(std::integral_constant<int, 1>{});
    break;
  case 2:
    by_steps(std::integral_constant<int, 2>{});
    break;
  case 3:
    by_steps(std::integral_constant<int, 3>{});
    break;
  default:
    by_steps(std::integral_constant<int, 4>{});
    break;
  }
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

template <typename S, int B, int VS>
void launch_rotate_long(trl_hip_ctx* ctx, int nt, size_t g, size_t shm, const S* v, size_t n, int m, const double* Y, int ybc, int j0, int nb, S* o, int ng2)
{
  switch (nt) {
  case 1:
    hipLaunchKernelGGL((k_rotate_long<S, B, 1, VS>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n, ng2);
    break;
  case 2:
    hipLaunchKernelGGL((k_rotate_long<S, B, 2, VS>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n, ng2);
    break;
  case 3:
    hipLaunchKernelGGL((k_rotate_long<S, B, 3, VS>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n, ng2);
    break;
  default:
    hipLaunchKernelGGL((k_rotate_long<S, B, 4, VS>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n, ng2);
    break;
  }
}

int launch_rotate(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int m, const void* V, const double* Y, int ybc, int kc, void* Out)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const int K = m * B;
      constexpr int VS = rot_vec<S, B>();
      const int nsup = (K + 4 * VS - 1) / (4 * VS);
      // a contraction that 16-byte super-steps would pad (K = 40 at B = 8 in
      // FP32: 48) goes in 8-byte super-steps that divide it (5 of 8)
      constexpr int VS2 = (VS == 4 && B % 2 == 0) ? 2 : VS;
      const bool exact2 = VS2 != VS && K % (4 * VS) != 0 && K % (4 * VS2) == 0 && K / (4 * VS2) <= kRotShortSteps;
      // buffer resources below 2^31 bytes: the Out blocks; V (short form), or
      // the blocks one super-step spans (long form)
      const size_t blk_bytes = n * B * sizeof(S);
      const size_t lim = size_t(1) << 31;
      const bool out_ok = !ctx->force_offsets64 && size_t(kc) * blk_bytes < lim;
      const bool short_k = out_ok && size_t(m) * blk_bytes < lim && (exact2 || nsup <= kRotShortSteps);
      // the long form's groups of kRotGroupK k-steps come in pairs: K padded to 2 kRotGroupK
      const int ng2 = 2 * ((K + 2 * kRotGroupK - 1) / (2 * kRotGroupK));
      const bool long_k = out_ok && size_t((4 * VS + B - 1) / B + 1) * blk_bytes < lim && !short_k;
      const int kpad = long_k ? ng2 * kRotGroupK : short_k ? 4 * (exact2 ? VS2 * (K / (4 * VS2)) : VS * nsup) : K;
      // output blocks per launch: NT*16 >= nb*B, LDS = kpad * NT*16 * sizeof(S) <= 144 KiB
      const size_t lds_cap = ctx->max_lds < size_t(144 * 1024) ? ctx->max_lds : size_t(144 * 1024);
      int nt_fit = int(lds_cap / (size_t(kpad) * 16 * sizeof(S)));
      if (nt_fit < 1) return TRL_HIP_EINVAL;
      if (nt_fit > 4) nt_fit = 4;
      int nb_max = (nt_fit * 16) / B;
      if (nb_max < 1) return TRL_HIP_EINVAL;
      size_t g = (n + 127) / 128;
      // 512-lane workgroups per CU, measured on the 160^3 stencil (b = 8):
      // the short form 2 in FP32 and 1 in FP64 (4: 287 against 272 us, 534
      // against 516 us), the long form 2 (restart, m = 32: 2.14 ms, 1 or 4:
      // 2.17-2.18 ms), the fallback form 4
      const size_t gmax = size_t(4) * ctx->num_cus;
      const size_t gmax_short = size_t(sizeof(S) == 4 ? 2 : 1) * ctx->num_cus, gmax_long = size_t(2) * ctx->num_cus;
      // fallback form, FP64 with many basis blocks: one element per load
      // (their register cost lowers the occupancy of its contraction loop)
      const bool v1 = sizeof(S) == 8 && m >= 8;
      const size_t g_short = g < gmax_short ? g : gmax_short, g_long = g < gmax_long ? g : gmax_long;
      if (g > gmax) g = gmax;
      for (int j0 = 0; j0 < kc; j0 += nb_max) {
        const int nb = (kc - j0) < nb_max ? (kc - j0) : nb_max;
        const int nt = (nb * B + 15) / 16;
        const size_t shm = size_t(K) * nt * 16 * sizeof(S);
        const S* v = static_cast<const S*>(V);
        S* o = static_cast<S*>(Out) + size_t(j0) * n * B;
        // short contractions: the whole operand row of a tile in registers, next tile prefetched
        if (short_k) {
          const int rc = exact2 ? launch_rotate_short<S, B, VS2>(ctx, nt, K / (4 * VS2), g_short, v, n, m, Y, ybc, j0, nb, o)
                                : launch_rotate_short<S, B, VS>(ctx, nt, nsup, g_short, v, n, m, Y, ybc, j0, nb, o);
          if (rc) return rc;
          continue;
        }
        // long contractions: groups of k-steps streamed through two register sets
        if (long_k) {
          launch_rotate_long<S, B, VS>(ctx, nt, g_long, size_t(ng2) * kRotGroupK * nt * 16 * sizeof(S), v, n, m, Y, ybc, j0, nb, o, ng2);
          TRL_HIP_LAUNCH_CHECK();
          continue;
        }
        switch (nt) {
        case 1:
          if (v1)
            hipLaunchKernelGGL((k_rotate<S, B, 1, 1>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          else
            hipLaunchKernelGGL((k_rotate<S, B, 1, rot_vec<S, B>()>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        case 2:
          if (v1)
            hipLaunchKernelGGL((k_rotate<S, B, 2, 1>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          else
            hipLaunchKernelGGL((k_rotate<S, B, 2, rot_vec<S, B>()>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        case 3:
          if (v1)
            hipLaunchKernelGGL((k_rotate<S, B, 3, 1>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          else
            hipLaunchKernelGGL((k_rotate<S, B, 3, rot_vec<S, B>()>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        default:
          if (v1)
            hipLaunchKernelGGL((k_rotate<S, B, 4, 1>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          else
            hipLaunchKernelGGL((k_rotate<S, B, 4, rot_vec<S, B>()>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        }
        TRL_HIP_LAUNCH_CHECK();
      }
    });
  });
  return 0;
}

} // namespace trl_hip_impl
