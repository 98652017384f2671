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

constexpr int kThreads = 256;
// FP32 Gram: rows folded into the compensated accumulator every kChunk quads
constexpr int kChunk = 8;

// ---------------------------------------------------------------------------
// Gram.  A tile stacks 16/BP blocks of V (columns padded to BP = pow2(b));
// B is diag(m) W padded to 16 columns; K runs over rows, 4 per instruction.
// A workgroup has 4 waves: TG tile groups x RS row subsets.  Wave (tg, rs)
// handles tiles t = tg + TG*q (q < TPW) on the row quads of subset rs and
// writes its partial sums to partials[(blockIdx.x*RS + rs)*P + idx], P = k*b*b.
template <typename S, int B, int TPW>
__global__ __launch_bounds__(kThreads) void k_gram(const S* __restrict__ V, size_t ldv, int k, const S* __restrict__ W, const S* __restrict__ md, size_t n,
                                                   int TG, double* __restrict__ partials)
{
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

  const size_t nquads = (n + 3) / 4;
  const size_t qstride = size_t(gridDim.x) * RS;
  size_t qd = size_t(blockIdx.x) * RS + rs;

  // FP64: one accumulator; FP32: chunk accumulator folded into (hi, lo) by TwoSum
  acc_t acc[TPW];
  acc_t lo[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    acc[q] = acc_t{0, 0, 0, 0};
    lo[q] = acc_t{0, 0, 0, 0};
  }

  if constexpr (sizeof(S) == 8) {
    for (; qd < nquads; qd += qstride) {
      const size_t r = qd * 4 + q4;
      const bool rok = r < n;
      S bw = 0;
      if (rok && wval) {
        bw = W[r * B + i16];
        if (md) bw *= md[r];
      }
#pragma unroll
      for (int q = 0; q < TPW; ++q) {
        S a = 0;
        if (rok && aval[q]) a = V[aoff[q] + r * B];
        acc[q] = M::op(a, bw, acc[q]);
      }
    }
  }
  else {
    while (qd < nquads) {
      acc_t cur[TPW];
#pragma unroll
      for (int q = 0; q < TPW; ++q) cur[q] = acc_t{0, 0, 0, 0};
      for (int it = 0; it < kChunk && qd < nquads; ++it, qd += qstride) {
        const size_t r = qd * 4 + q4;
        const bool rok = r < n;
        S bw = 0;
        if (rok && wval) {
          bw = W[r * B + i16];
          if (md) bw *= md[r];
        }
#pragma unroll
        for (int q = 0; q < TPW; ++q) {
          S a = 0;
          if (rok && aval[q]) a = V[aoff[q] + r * B];
          cur[q] = M::op(a, bw, cur[q]);
        }
      }
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

  const size_t P = size_t(k) * B * B;
  double* out = partials + (size_t(blockIdx.x) * RS + rs) * P;
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
        out[(size_t(blk) * B + ci) * B + i16] = v;
      }
    }
  }
}

// out[p] = sum_{w < nw} partials[w*P + p], summed in fixed order (deterministic).
__global__ __launch_bounds__(kThreads) void k_reduce(const double* __restrict__ partials, int nw, int P, double* __restrict__ out)
{
  __shared__ double red[4][64];
  const int pl = threadIdx.x & 63, s = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + pl;
  double acc = 0.0;
  if (p < P)
    for (int w = s; w < nw; w += 4) acc += partials[size_t(w) * P + p];
  red[s][pl] = acc;
  __syncthreads();
  if (s == 0 && p < P) out[p] = (red[0][pl] + red[1][pl]) + (red[2][pl] + red[3][pl]);
}

// Sum the b x b Gram partials, factor G = R^T R in FP64 (lower triangle of G
// read, as an LLT does), write R (times Rprev if given) and R^{-1}.  On
// breakdown the sticky flag receives the 1-based pivot, R and R^{-1} are set
// to the identity so the following scale leaves W unchanged.
__global__ __launch_bounds__(kThreads) void k_chol(const double* __restrict__ partials, int nw, int b, const double* __restrict__ Rprev,
                                                   double* __restrict__ Rout, double* __restrict__ Rinv, int* __restrict__ flags)
{
  __shared__ double red[4][64];
  __shared__ double G[256];
  const int P = b * b;
  for (int base = 0; base < P; base += 64) {
    const int pl = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int p = base + pl;
    double acc = 0.0;
    if (p < P)
      for (int w = s; w < nw; w += 4) acc += partials[size_t(w) * P + p];
    red[s][pl] = acc;
    __syncthreads();
    if (s == 0 && p < P) G[p] = (red[0][pl] + red[1][pl]) + (red[2][pl] + red[3][pl]);
    __syncthreads();
  }
  __shared__ double Rl[16][16];
  __shared__ double Xl[16][16];
  if (threadIdx.x != 0) return;
  auto& R = Rl;
  auto& X = Xl;
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
  // X = R^{-1} (upper triangular), column by column
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
  if (Rprev) {
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < b; ++j) {
        double s = 0.0;
        for (int p = i; p <= j; ++p) s += R[i][p] * Rprev[p * b + j];
        Rout[i * b + j] = (j >= i) ? s : 0.0;
      }
  }
  else {
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < b; ++j) Rout[i * b + j] = R[i][j];
  }
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
__global__ __launch_bounds__(kThreads) void k_update(const S* __restrict__ V, size_t ldv, int k, const double* __restrict__ C, S* __restrict__ W, size_t n)
{
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
template <typename S, int B, int NT>
__global__ __launch_bounds__(512) void k_rotate(const S* __restrict__ V, size_t ldv, int m, const double* __restrict__ Yg, int ybc, int jb0, int nb,
                                                S* __restrict__ Out, size_t n)
{
  using M = Mfma<S>;
  using acc_t = typename M::acc_t;
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
  const int ksteps = (K + 3) / 4;
  for (size_t rt = size_t(blockIdx.x) * nwave + wave; rt < ntiles; rt += size_t(gridDim.x) * nwave) {
    const size_t r0 = rt * 16;
    const size_t ra = r0 + i16;
    const bool rok = ra < n;
    acc_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = acc_t{0, 0, 0, 0};
    for (int s = 0; s < ksteps; ++s) {
      const int kk = 4 * s + q4;
      const bool kok = kk < K;
      S a = 0;
      if (rok && kok) {
        const int j = kk / B, c = kk % B;
        a = V[size_t(j) * ldv + ra * B + c];
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const S bv = kok ? Ys[kk * ldy + t * 16 + i16] : S(0);
        acc[t] = M::op(a, bv, acc[t]);
      }
    }
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

inline int pick_tpw(int per_wave)
{
  if (per_wave <= 1) return 1;
  if (per_wave <= 2) return 2;
  if (per_wave <= 4) return 4;
  if (per_wave <= 6) return 6;
  if (per_wave <= 8) return 8;
  if (per_wave <= 12) return 12;
  return 16;
}

// Queue the Gram partials for blocks [0, k) of V against W; returns the
// number of partial rows written (nwg * RS) in *nw.
template <typename S, int B>
int queue_gram_partials(trl_hip_ctx* ctx, size_t n, int k, const S* V, const S* W, const S* md, int* nw)
{
  constexpr int BP = pad_b(B), BPT = 16 / BP;
  const int ntile = (k + BPT - 1) / BPT;
  const int TG = ntile >= 4 ? 4 : (ntile >= 2 ? 2 : 1);
  const int RS = 4 / TG;
  const int tpw = pick_tpw((ntile + TG - 1) / TG);
  if ((ntile + TG - 1) / TG > 16) return TRL_HIP_EINVAL;
  const size_t nquads = (n + 3) / 4;
  size_t g = (nquads + RS - 1) / RS;
  const size_t gmax = size_t(2) * ctx->num_cus;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  const size_t P = size_t(k) * B * B;
  int rc = trl_hip_impl::ensure_partials(ctx, g * RS * P);
  if (rc) return rc;
  const size_t ldv = n * B;
  const dim3 grid{unsigned(g)}, blk{unsigned(kThreads)};
  switch (tpw) {
  case 1:
    hipLaunchKernelGGL((k_gram<S, B, 1>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, ctx->partials);
    break;
  case 2:
    hipLaunchKernelGGL((k_gram<S, B, 2>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, ctx->partials);
    break;
  case 4:
    hipLaunchKernelGGL((k_gram<S, B, 4>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, ctx->partials);
    break;
  case 6:
    hipLaunchKernelGGL((k_gram<S, B, 6>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, ctx->partials);
    break;
  case 8:
    hipLaunchKernelGGL((k_gram<S, B, 8>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, ctx->partials);
    break;
  case 12:
    hipLaunchKernelGGL((k_gram<S, B, 12>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, ctx->partials);
    break;
  default:
    hipLaunchKernelGGL((k_gram<S, B, 16>), grid, blk, 0, ctx->stream, V, ldv, k, W, md, n, TG, ctx->partials);
    break;
  }
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

int launch_gram(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const void* W, const void* md, double* C)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const S* v = static_cast<const S*>(V);
      const S* w = static_cast<const S*>(W);
      const S* m = static_cast<const S*>(md);
      if (n == 0) {
        hipLaunchKernelGGL(k_reduce, dim3(unsigned((size_t(k) * B * B + 63) / 64)), dim3(kThreads), 0, ctx->stream, ctx->partials, 0, k * B * B, C);
        TRL_HIP_LAUNCH_CHECK();
        return 0;
      }
      constexpr int kmax = gram_max_k<B>();
      for (int j0 = 0; j0 < k; j0 += kmax) {
        const int kc = (k - j0) < kmax ? (k - j0) : kmax;
        int nw = 0;
        int rc = queue_gram_partials<S, B>(ctx, n, kc, v + size_t(j0) * n * B, w, m, &nw);
        if (rc) return rc;
        const int P = kc * B * B;
        hipLaunchKernelGGL(k_reduce, dim3(unsigned((P + 63) / 64)), dim3(kThreads), 0, ctx->stream, ctx->partials, nw, P, C + size_t(j0) * B * B);
        TRL_HIP_LAUNCH_CHECK();
      }
    });
  });
  return 0;
}

int launch_update(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const double* C, void* W)
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
                           C + size_t(j0) * B * B, static_cast<S*>(W), n);
        TRL_HIP_LAUNCH_CHECK();
      }
    });
  });
  return 0;
}

int launch_cholqr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, const void* md, double* R, int passes)
{
  double* Rinv = ctx->small;
  double* R1 = ctx->small + 256;
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      S* w = static_cast<S*>(W);
      const S* m = static_cast<const S*>(md);
      for (int pass = 0; pass < passes; ++pass) {
        int nw = 0;
        if (n > 0) {
          int rc = queue_gram_partials<S, B>(ctx, n, 1, w, w, m, &nw);
          if (rc) return rc;
        }
        const bool last = (pass == passes - 1);
        double* rout = last ? R : R1;
        const double* rprev = (pass > 0) ? R1 : nullptr;
        hipLaunchKernelGGL(k_chol, dim3(1), dim3(kThreads), 0, ctx->stream, ctx->partials, nw, B, rprev, rout, Rinv, ctx->flags);
        TRL_HIP_LAUNCH_CHECK();
        if (n > 0) {
          size_t g = (n + kThreads - 1) / kThreads;
          const size_t gmax = size_t(8) * ctx->num_cus;
          if (g > gmax) g = gmax;
          hipLaunchKernelGGL((k_scale<S, B>), dim3(unsigned(g)), dim3(kThreads), 0, ctx->stream, w, Rinv, n);
          TRL_HIP_LAUNCH_CHECK();
        }
      }
    });
  });
  return 0;
}

int launch_rotate(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int m, const void* V, const double* Y, int ybc, int kc, void* Out)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const int K = m * B;
      // output blocks per launch: NT*16 >= nb*B, LDS = K * NT*16 * sizeof(S) <= 144 KiB
      const size_t lds_cap = ctx->max_lds < size_t(144 * 1024) ? ctx->max_lds : size_t(144 * 1024);
      int nt_fit = int(lds_cap / (size_t(K) * 16 * sizeof(S)));
      if (nt_fit < 1) return TRL_HIP_EINVAL;
      if (nt_fit > 4) nt_fit = 4;
      int nb_max = (nt_fit * 16) / B;
      if (nb_max < 1) return TRL_HIP_EINVAL;
      size_t g = (n + 127) / 128;
      const size_t gmax = size_t(2) * ctx->num_cus;
      if (g > gmax) g = gmax;
      for (int j0 = 0; j0 < kc; j0 += nb_max) {
        const int nb = (kc - j0) < nb_max ? (kc - j0) : nb_max;
        const int nt = (nb * B + 15) / 16;
        const size_t shm = size_t(K) * nt * 16 * sizeof(S);
        const S* v = static_cast<const S*>(V);
        S* o = static_cast<S*>(Out) + size_t(j0) * n * B;
        switch (nt) {
        case 1:
          hipLaunchKernelGGL((k_rotate<S, B, 1>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        case 2:
          hipLaunchKernelGGL((k_rotate<S, B, 2>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        case 3:
          hipLaunchKernelGGL((k_rotate<S, B, 3>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        default:
          hipLaunchKernelGGL((k_rotate<S, B, 4>), dim3(unsigned(g)), dim3(512), shm, ctx->stream, v, n * B, m, Y, ybc, j0, nb, o, n);
          break;
        }
        TRL_HIP_LAUNCH_CHECK();
      }
    });
  });
  return 0;
}

} // namespace trl_hip_impl
