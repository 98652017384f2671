// Batched block Lanczos: many independent sparse eigenproblems advanced in
// lock-step, one persistent kernel launch per Lanczos step.
//
// Layout.  The problems' rows are concatenated (problem p owns rows
// [roff_p, roff_p + n_p) of an N-row index space) and their matrices form one
// block-diagonal operator with absolute column indices (CSR, or slot-major
// ELL over N rows).  Basis block q of every problem is one N x b row-major
// panel, so block q of problem p is the contiguous sub-panel starting at row
// roff_p.  The rows are cut into tiles that never straddle two problems; one
// workgroup owns one tile for a whole step.
//
// One step (the extend loop of include/trl/eigensolvers/lanczos.hh:244-284,
// or the quasi-Lanczos step after a restart, lanczos.hh:150-178) for every
// active problem is ONE launch of k_batch_step: each workgroup keeps its
// rows of the new block in registers through all phases
//   1  W = A V_i (- V_{i-1} beta^T)             partial alpha = V_i^T W
//   2  W -= V_i alpha (quasi: - V_{i-1} T_{i,i-1}^T), CGS pass 1 partials
//   3  W -= V C1, CGS pass 2 partials
//   4  W -= V C2, Gram partials
//   5  FP64 Cholesky (every workgroup of the problem, same order), W R^{-1}
// and synchronises with the other workgroups of its problem through a
// device-side barrier (an arrival counter per problem and launch) after each
// partial.  Reductions of the partials run in a fixed order, so every
// workgroup of a problem derives bit-identical coefficients and the result is
// deterministic.  Barriers spin at most kMaxSpins times: on timeout they set
// an error flag that every workgroup checks, so every launch terminates.  The
// host only launches grids that fit on the device at once (occupancy check).
//
// The steps of one restart cycle are captured once into a hipGraph (a fill
// kernel for the counters, then one k_batch_step per step) and replayed every
// cycle; the kernels read the active set and the basis selector from device
// memory, so the same graph serves every cycle.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "common.hh"

using namespace trl_hip_dev;

namespace {

constexpr int kT = 256;
constexpr int kChunkB = 8;
constexpr int kMaxB = 8;
constexpr int kBB = kMaxB * kMaxB;
constexpr unsigned kMaxSpins = 1u << 22;

struct BatchDev {
  int ntiles;
  int m;
  int slot; // doubles per tile in one partial buffer ((m + 1) * B * B)
  int nseq; // counter slots per problem (one per launch in a cycle)
  size_t N;
  const int* tile_prob;
  const int* tile_row0;
  const int* tile_row1;
  const int* tile_local;
  const int* prob_ntiles;
  const int* prob_tile0;
  const int* active;
  const int* vsel; // 0: basis in V0, 1: in V1
  const int32_t* rowptr; // NULL: ELL (width slots, slot-major over N rows)
  const int32_t* col;
  const double* val;
  int width;
  double* V0;
  double* V1;
  double* partials; // [2][ntiles][slot]
  unsigned* counters; // [nprob][nseq]
  double* alpha; // [nprob][m][B*B]
  double* beta; // [nprob][m][B*B]
  double* beta_prev; // [nprob][B*B]
  double* arrow; // [nprob][B*B]
  int* pflag; // [nprob] Cholesky breakdown pivot (1-based), sticky
  int* err; // barrier timeout
};

__device__ inline int load_flag(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Arrive at and wait for `target` arrivals on *cnt (all tiles of one
// problem).  Releases this workgroup's writes before arriving and acquires
// the others' after the wait.  Returns false if a timeout was raised.
__device__ bool tile_barrier(unsigned* cnt, unsigned target, int* err)
{
  __shared__ int s_ok;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(cnt, 1u);
    int ok = 1;
    unsigned spins = 0;
    while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (load_flag(err)) {
        ok = 0;
        break;
      }
      if (++spins > kMaxSpins) {
        atomicExch(err, 1);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Sum E per-thread values over the workgroup (butterfly per wave, then the
// four waves in order) and store them at dst (global).
template <int E>
__device__ inline void wg_reduce_store(const double (&acc)[E], double* dst, double* red)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    double v = acc[e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave * E + e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < E; e += kT) dst[e] = ((red[e] + red[E + e]) + red[2 * E + e]) + red[3 * E + e];
  __syncthreads();
}

// out[e] = sum over the problem's T tiles of part[tile][e], e < E, in a fixed
// order: S = kT / Ec lanes per entry stride over the tiles, then the lanes'
// sums are added in lane order.
__device__ inline void reduce_partials(const double* part, int slot, int tile0, int T, int E, double* out, double* scratch)
{
  // at least 4 lanes per entry, each with 4 loads in flight: the partials
  // of a large problem are summed in a few memory latencies, not T
  constexpr int kMinLanes = 4;
  for (int e0 = 0; e0 < E; e0 += kT / kMinLanes) {
    const int Ec = min(kT / kMinLanes, E - e0);
    const int S = kT / Ec;
    const int e = threadIdx.x % Ec, s = threadIdx.x / Ec;
    double acc = 0.0;
    if (s < S) {
      const double* pe = part + size_t(tile0) * slot + e0 + e;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int t = s;
      for (; t + 3 * S < T; t += 4 * S) {
        a0 += pe[size_t(t) * slot];
        a1 += pe[size_t(t + S) * slot];
        a2 += pe[size_t(t + 2 * S) * slot];
        a3 += pe[size_t(t + 3 * S) * slot];
      }
      for (; t < T; t += S) a0 += pe[size_t(t) * slot];
      acc = (a0 + a1) + (a2 + a3);
    }
    scratch[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < Ec) {
      double v = 0.0;
      for (int q = 0; q < S; ++q) v += scratch[q * Ec + threadIdx.x];
      out[e0 + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// G (B x B, symmetric) = R^T R; Rinv = R^{-1}.  On breakdown returns the
// 1-based pivot and sets R = Rinv = I.  One thread; L, X: B*B work arrays
// (LDS).
template <int B>
__device__ int chol_small(const double* G, double* R, double* Rinv, double* L, double* X)
{
  for (int e = 0; e < B * B; ++e) L[e] = X[e] = 0.0;
  int fail = 0;
  for (int j = 0; j < B && !fail; ++j) {
    double dd = G[j * B + j];
    for (int q = 0; q < j; ++q) dd -= L[q * B + j] * L[q * B + j];
    if (!(dd > 0.0)) {
      fail = j + 1;
      break;
    }
    const double rjj = sqrt(dd);
    L[j * B + j] = rjj;
    for (int i = j + 1; i < B; ++i) {
      double s = G[i * B + j];
      for (int q = 0; q < j; ++q) s -= L[q * B + j] * L[q * B + i];
      L[j * B + i] = s / rjj;
    }
  }
  if (fail) {
    for (int e = 0; e < B * B; ++e) R[e] = Rinv[e] = (e % (B + 1) == 0) ? 1.0 : 0.0;
    return fail;
  }
  for (int j = 0; j < B; ++j) {
    X[j * B + j] = 1.0 / L[j * B + j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int q = i + 1; q <= j; ++q) s += L[i * B + q] * X[q * B + j];
      X[i * B + j] = -s / L[i * B + i];
    }
  }
  for (int e = 0; e < B * B; ++e) {
    R[e] = L[e];
    Rinv[e] = X[e];
  }
  return 0;
}

template <int B, bool ELL>
__device__ inline void spmm_row(const BatchDev& d, const double* X, int r, double (&y)[B])
{
  constexpr int kCh = B >= 4 ? 4 : kChunkB; // nonzeros whose gathers are in flight together
#pragma unroll
  for (int c = 0; c < B; ++c) y[c] = 0.0;
  const int beg = ELL ? 0 : d.rowptr[r];
  const int end = ELL ? d.width : d.rowptr[r + 1];
  for (int p0 = beg; p0 < end; p0 += kCh) {
    int ci[kCh];
    double vi[kCh];
#pragma unroll
    for (int q = 0; q < kCh; ++q) {
      const bool ok = p0 + q < end;
      const size_t p = ELL ? size_t(p0 + q) * d.N + size_t(r) : size_t(p0 + q);
      ci[q] = ok ? d.col[p] : r;
      vi[q] = ok ? d.val[p] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kCh; ++q) {
      double x[B];
      RowIO<double, B>::load(X + size_t(ci[q]) * B, x);
#pragma unroll
      for (int c = 0; c < B; ++c) y[c] = fma(vi[q], x[c], y[c]);
    }
  }
}

// acc[c*B + d] += x[c] * y[d]
template <int B>
__device__ inline void outer_acc(const double (&x)[B], const double (&y)[B], double (&acc)[B * B])
{
#pragma unroll
  for (int c = 0; c < B; ++c)
#pragma unroll
    for (int dd = 0; dd < B; ++dd) acc[c * B + dd] = fma(x[c], y[dd], acc[c * B + dd]);
}

// y[d] -= sum_c x[c] * M[c*B + d]   (y -= x M)
template <int B>
__device__ inline void sub_xm(const double (&x)[B], const double* M, double (&y)[B])
{
#pragma unroll
  for (int dd = 0; dd < B; ++dd) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < B; ++c) s = fma(x[c], M[c * B + dd], s);
    y[dd] -= s;
  }
}

// y[c] -= sum_d x[d] * M[c*B + d]   (y -= x M^T)
template <int B>
__device__ inline void sub_xmt(const double (&x)[B], const double* M, double (&y)[B])
{
#pragma unroll
  for (int c = 0; c < B; ++c) {
    double s = 0.0;
#pragma unroll
    for (int dd = 0; dd < B; ++dd) s = fma(x[dd], M[c * B + dd], s);
    y[c] -= s;
  }
}

// kind 0: extend step i; 1: quasi-Lanczos step at i = k after a restart;
// 2: orthonormalise block 0 only.  seq: index of this launch in its cycle
// (selects the barrier counter).  The workgroup's rows of the new block W
// live in the output block itself between phases (written and re-read by
// the same workgroup, so they stay in its caches); the threads stride over
// the tile's rows.  Dynamic LDS: (m+1)*B*B doubles (CGS coefficients, and
// W^T W for the second-pass test).  adaptive: the second Gram-Schmidt pass
// runs only for the problems whose first pass cancelled a column below
// 1/sqrt(2) of its norm (the DGKS test of the single-problem path), decided
// identically by every workgroup of a problem from the reduced coefficients.
template <int B, bool ELL>
__global__ __launch_bounds__(kT, B >= 8 ? 1 : 4) void k_batch_step(BatchDev d, int i, int kind, int seq, int cq_passes, int adaptive)
{
  constexpr int BB = B * B;
  extern __shared__ double Cs[];
  __shared__ double red[4 * BB];
  __shared__ double scratch[kT];
  __shared__ double al[BB], Gs[BB], Rp[BB], Ri[BB], Racc[BB], Rt[BB], aux[BB], Lw[BB], Xw[BB];
  __shared__ int fail_s;

  const int t = blockIdx.x;
  if (t >= d.ntiles) return;
  const int p = d.tile_prob[t];
  if (!d.active[p] || load_flag(d.err)) return;
  const int T = d.prob_ntiles[p], tile0 = d.prob_tile0[p], lt = d.tile_local[t];
  const int r0 = d.tile_row0[t], r1 = d.tile_row1[t];
  const int rb = r0 + int(threadIdx.x);
  unsigned* cnt = d.counters + size_t(p) * d.nseq + seq;
  unsigned arrived = 0;
  int buf = 0;
  const size_t NB = d.N * B;
  double* V = *d.vsel ? d.V1 : d.V0;
  double* W = V + size_t(kind == 2 ? 0 : i + 1) * NB; // output block (holds W between phases)
  auto part = [&](int b_) { return d.partials + size_t(b_) * d.ntiles * d.slot; };
  auto barrier = [&]() {
    ++arrived;
    return tile_barrier(cnt, arrived * unsigned(T), d.err);
  };
  auto row = [&](const double* X, int r, double (&v)[B]) { RowIO<double, B>::load(X + size_t(r) * B, v); };
  auto put = [&](int r, const double (&v)[B]) { RowIO<double, B>::store(W + size_t(r) * B, v); };

  double acc[BB];
  if (kind != 2) {
    const double* Vc = V + size_t(i) * NB;
    const double* Vp = i > 0 ? V + size_t(i - 1) * NB : nullptr;
    const bool sub_beta = (kind == 0 && i > 0);
    if (threadIdx.x < BB) aux[threadIdx.x] = sub_beta ? d.beta_prev[size_t(p) * BB + threadIdx.x] : (kind == 1 ? d.arrow[size_t(p) * BB + threadIdx.x] : 0.0);
    __syncthreads();
    // ---- phase 1: W = A V_i (- V_{i-1} beta^T); alpha partial = V_i^T W ----
#pragma unroll
    for (int e = 0; e < BB; ++e) acc[e] = 0.0;
    for (int r = rb; r < r1; r += kT) {
      double y[B], vc[B];
      spmm_row<B, ELL>(d, Vc, r, y);
      if (sub_beta) {
        double vp[B];
        row(Vp, r, vp);
        sub_xmt<B>(vp, aux, y);
      }
      row(Vc, r, vc);
      outer_acc<B>(vc, y, acc);
      put(r, y);
    }
    wg_reduce_store<BB>(acc, part(buf) + size_t(t) * d.slot, red);
    if (!barrier()) return;
    reduce_partials(part(buf), d.slot, tile0, T, BB, al, scratch);
    buf ^= 1;
    // ---- phase 2: W -= V_i alpha (quasi: W -= V_{i-1} T_{i,i-1}^T) ----------
    for (int r = rb; r < r1; r += kT) {
      double y[B], vc[B];
      row(W, r, y);
      row(Vc, r, vc);
      sub_xm<B>(vc, al, y);
      if (kind == 1 && Vp) {
        double vp[B];
        row(Vp, r, vp);
        sub_xmt<B>(vp, aux, y);
      }
      put(r, y);
    }
    // ---- classical Gram-Schmidt against blocks 0..i: two passes, or one
    // when the test after the first finds no cancellation ------------------
    const int nq = i + 1;
    __shared__ int need_s;
    for (int pass = 0; pass < 2; ++pass) {
      const bool test = adaptive && pass == 0;
      for (int q = 0; q < nq + (test ? 1 : 0); ++q) {
        const double* Vq = q < nq ? V + size_t(q) * NB : nullptr; // q == nq: W^T W
#pragma unroll
        for (int e = 0; e < BB; ++e) acc[e] = 0.0;
        for (int r = rb; r < r1; r += kT) {
          double y[B], vq[B];
          row(W, r, y);
          if (Vq)
            row(Vq, r, vq);
          else
#pragma unroll
            for (int c = 0; c < B; ++c) vq[c] = y[c];
          outer_acc<B>(vq, y, acc);
        }
        wg_reduce_store<BB>(acc, part(buf) + size_t(t) * d.slot + size_t(q) * BB, red);
      }
      if (!barrier()) return;
      reduce_partials(part(buf), d.slot, tile0, T, (nq + (test ? 1 : 0)) * BB, Cs, scratch);
      buf ^= 1;
      if (test) {
        // column j: ||w_j||^2 after the pass = ||w_j||^2 - sum_q ||c_qj||^2
        if (threadIdx.x == 0) need_s = 0;
        __syncthreads();
        if (threadIdx.x < unsigned(B)) {
          const int j = int(threadIdx.x);
          const double w2 = Cs[size_t(nq) * BB + j * B + j];
          double c2 = 0.0;
          for (int q = 0; q < nq; ++q)
            for (int c = 0; c < B; ++c) {
              const double x = Cs[size_t(q) * BB + c * B + j];
              c2 = fma(x, x, c2);
            }
          if (!(w2 - c2 >= 0.5 * w2) || !(w2 > 0.0)) atomicOr(&need_s, 1);
        }
        __syncthreads();
      }
      for (int r = rb; r < r1; r += kT) {
        double y[B];
        row(W, r, y);
        for (int q = 0; q < nq; ++q) {
          double vq[B];
          row(V + size_t(q) * NB, r, vq);
          sub_xm<B>(vq, Cs + q * BB, y);
        }
        put(r, y);
      }
      if (test && !need_s) break;
    }
  }
  // ---- Cholesky-QR of W (cq_passes passes, R = R_2 R_1) ---------------------
  if (threadIdx.x < BB) Racc[threadIdx.x] = (threadIdx.x % (B + 1) == 0) ? 1.0 : 0.0;
  if (threadIdx.x == 0) fail_s = 0;
  __syncthreads();
  for (int pass = 0; pass < cq_passes; ++pass) {
#pragma unroll
    for (int e = 0; e < BB; ++e) acc[e] = 0.0;
    for (int r = rb; r < r1; r += kT) {
      double y[B];
      row(W, r, y);
      outer_acc<B>(y, y, acc);
    }
    wg_reduce_store<BB>(acc, part(buf) + size_t(t) * d.slot, red);
    if (!barrier()) return;
    reduce_partials(part(buf), d.slot, tile0, T, BB, Gs, scratch);
    buf ^= 1;
    if (threadIdx.x == 0) {
      const int f = chol_small<B>(Gs, Rp, Ri, Lw, Xw);
      if (f && !fail_s) fail_s = f;
      for (int a = 0; a < B; ++a) // Racc <- Rp Racc
        for (int c = 0; c < B; ++c) {
          double sum = 0.0;
          for (int q = 0; q < B; ++q) sum += Rp[a * B + q] * Racc[q * B + c];
          Rt[a * B + c] = sum;
        }
      for (int e = 0; e < BB; ++e) Racc[e] = Rt[e];
    }
    __syncthreads();
    for (int r = rb; r < r1; r += kT) {
      double y[B], o[B];
      row(W, r, y);
#pragma unroll
      for (int dd = 0; dd < B; ++dd) {
        double sum = 0.0;
#pragma unroll
        for (int c = 0; c <= dd; ++c) sum = fma(y[c], Ri[c * B + dd], sum);
        o[dd] = sum;
      }
      put(r, o);
    }
  }
  if (lt == 0) {
    if (threadIdx.x == 0 && fail_s && d.pflag[p] == 0) d.pflag[p] = fail_s;
    if (kind != 2 && threadIdx.x < BB) {
      d.alpha[(size_t(p) * d.m + i) * BB + threadIdx.x] = al[threadIdx.x];
      d.beta[(size_t(p) * d.m + i) * BB + threadIdx.x] = Racc[threadIdx.x];
      d.beta_prev[size_t(p) * BB + threadIdx.x] = Racc[threadIdx.x];
    }
  }
}

__global__ void k_zero_u32(unsigned* p, size_t n)
{
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0u;
}

// Thick-restart rotation for every active problem:
// Out_j = sum_{q<m} V_q Y_p(q, j) for j < kr, Out_kr = V_m (inactive
// problems copy their blocks 0..kr unchanged).  Y_p is
// (m B) x (kr B) row-major at Y + p*(m B)*(kr B).  Dynamic LDS: Y_p.
template <int B>
__global__ __launch_bounds__(kT) void k_batch_rotate(BatchDev d, const double* __restrict__ Y, int kr)
{
  extern __shared__ double Ys[];
  const int t = blockIdx.x;
  if (t >= d.ntiles) return;
  const int p = d.tile_prob[t];
  const size_t NB = d.N * B;
  const double* V = *d.vsel ? d.V1 : d.V0;
  double* O = *d.vsel ? d.V0 : d.V1;
  if (!d.active[p]) {
    // a finished problem keeps its leading blocks: carry them to the other buffer
    for (int q = 0; q <= kr; ++q)
      for (int r = d.tile_row0[t] + int(threadIdx.x); r < d.tile_row1[t]; r += kT) {
        double v[B];
        RowIO<double, B>::load(V + size_t(q) * NB + size_t(r) * B, v);
        RowIO<double, B>::store(O + size_t(q) * NB + size_t(r) * B, v);
      }
    return;
  }
  const int m = d.m, ldy = kr * B;
  const size_t ysz = size_t(m) * B * ldy;
  for (size_t e = threadIdx.x; e < ysz; e += kT) Ys[e] = Y[size_t(p) * ysz + e];
  __syncthreads();
  for (int r = d.tile_row0[t] + int(threadIdx.x); r < d.tile_row1[t]; r += kT) {
    for (int j = 0; j < kr; ++j) {
      double o[B];
#pragma unroll
      for (int c = 0; c < B; ++c) o[c] = 0.0;
      for (int q = 0; q < m; ++q) {
        double v[B];
        RowIO<double, B>::load(V + size_t(q) * NB + size_t(r) * B, v);
#pragma unroll
        for (int a = 0; a < B; ++a)
#pragma unroll
          for (int c = 0; c < B; ++c) o[c] = fma(v[a], Ys[size_t(q * B + a) * ldy + j * B + c], o[c]);
      }
      RowIO<double, B>::store(O + size_t(j) * NB + size_t(r) * B, o);
    }
    double v[B];
    RowIO<double, B>::load(V + size_t(m) * NB + size_t(r) * B, v);
    RowIO<double, B>::store(O + size_t(kr) * NB + size_t(r) * B, v);
  }
}

// Residual sums of the Ritz pairs held in blocks 0..nblk-1 (standard
// problems): per tile and block, ||A x_c - lambda_c x_c||^2 and ||A x_c||^2
// for every column c, stored at out[(t*nblk + q)*2B ...].
template <int B, bool ELL>
__global__ __launch_bounds__(kT) void k_batch_residual(BatchDev d, const double* __restrict__ V, int nblk, const double* __restrict__ lam,
                                                        double* __restrict__ out)
{
  __shared__ double red[4 * 2 * B];
  const int t = blockIdx.x;
  if (t >= d.ntiles) return;
  const int p = d.tile_prob[t];
  const size_t NB = d.N * B;
  for (int q = 0; q < nblk; ++q) {
    double acc[2 * B];
#pragma unroll
    for (int e = 0; e < 2 * B; ++e) acc[e] = 0.0;
    const double* X = V + size_t(q) * NB;
    for (int r = d.tile_row0[t] + int(threadIdx.x); r < d.tile_row1[t]; r += kT) {
      double ax[B], x[B];
      spmm_row<B, ELL>(d, X, r, ax);
      RowIO<double, B>::load(X + size_t(r) * B, x);
#pragma unroll
      for (int c = 0; c < B; ++c) {
        const double rr = ax[c] - lam[(size_t(p) * nblk + q) * B + c] * x[c];
        acc[c] = fma(rr, rr, acc[c]);
        acc[B + c] = fma(ax[c], ax[c], acc[B + c]);
      }
    }
    wg_reduce_store<2 * B>(acc, out + (size_t(t) * nblk + q) * 2 * B, red);
  }
}

} // namespace

// ---------------------------------------------------------------------------
// Plan: device layout, tiles, graphs.

struct trl_hip_batch {
  trl_hip_ctx* ctx = nullptr;
  int nprob = 0, b = 0, m = 0;
  size_t N = 0;
  std::vector<size_t> n, roff;
  std::vector<int> active_h;
  // operator (absolute column indices)
  int32_t* rowptr = nullptr;
  int32_t* col = nullptr;
  double* val = nullptr;
  int width = 0;
  int32_t* ell_col = nullptr;
  double* ell_val = nullptr;
  long long nnz = 0;
  // variant
  int format = 0, rpt = 4;
  double tuned_us = 0.0;
  // tiles
  int ntiles = 0;
  std::vector<int> tile_prob_h, tile_row0_h, tile_row1_h, prob_ntiles_h, prob_tile0_h;
  int *tile_prob = nullptr, *tile_row0 = nullptr, *tile_row1 = nullptr, *tile_local = nullptr, *prob_ntiles = nullptr, *prob_tile0 = nullptr;
  // state
  double* V[2] = {nullptr, nullptr};
  int vsel_h = 0;
  int* vsel = nullptr;
  int* active = nullptr;
  double* partials = nullptr;
  size_t partials_cap = 0;
  unsigned* counters = nullptr;
  double *alpha = nullptr, *beta = nullptr, *beta_prev = nullptr, *arrow = nullptr, *Ybuf = nullptr, *lam = nullptr, *rsum = nullptr;
  size_t ybuf_count = 0; // doubles allocated at Ybuf
  size_t rsum_cap = 0, lam_cap = 0;
  int* pflag = nullptr;
  int* err = nullptr;
  int cq_passes = 1;
  int adaptive = 0; // second Gram-Schmidt pass only where the DGKS test asks
  bool use_graph = true;
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
  // leading blocks of finished problems (nblk_saved blocks of N x b), kept
  // out of the basis buffers so that the tiles need only cover active problems
  double* Xres = nullptr;
  int nblk_saved = 0;
  std::vector<int> saved_h;
};

namespace {

using trl_hip_impl::valid_b;

template <class T>
int dalloc(T** p, size_t count)
{
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return TRL_HIP_ENOMEM;
  *p = static_cast<T*>(q);
  return 0;
}

template <class T>
void dfree(T*& p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

void clear_graphs(trl_hip_batch* bt)
{
  for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
  bt->graphs.clear();
}

BatchDev make_dev(trl_hip_batch* bt)
{
  BatchDev d{};
  d.ntiles = bt->ntiles;
  d.m = bt->m;
  d.slot = (bt->m + 1) * bt->b * bt->b;
  d.nseq = bt->m + 1;
  d.N = bt->N;
  d.tile_prob = bt->tile_prob;
  d.tile_row0 = bt->tile_row0;
  d.tile_row1 = bt->tile_row1;
  d.tile_local = bt->tile_local;
  d.prob_ntiles = bt->prob_ntiles;
  d.prob_tile0 = bt->prob_tile0;
  d.active = bt->active;
  d.vsel = bt->vsel;
  if (bt->format == 1) {
    d.rowptr = nullptr;
    d.col = bt->ell_col;
    d.val = bt->ell_val;
    d.width = bt->width;
  }
  else {
    d.rowptr = bt->rowptr;
    d.col = bt->col;
    d.val = bt->val;
    d.width = 0;
  }
  d.V0 = bt->V[0];
  d.V1 = bt->V[1];
  d.partials = bt->partials;
  d.counters = bt->counters;
  d.alpha = bt->alpha;
  d.beta = bt->beta;
  d.beta_prev = bt->beta_prev;
  d.arrow = bt->arrow;
  d.pflag = bt->pflag;
  d.err = bt->err;
  return d;
}

// dispatch on (b, format); f receives the kernel function pointer
template <class F>
int with_step_kernel(int b, int format, F&& f)
{
  switch (b) {
  case 1:
    return format == 1 ? f(&k_batch_step<1, true>) : f(&k_batch_step<1, false>);
  case 2:
    return format == 1 ? f(&k_batch_step<2, true>) : f(&k_batch_step<2, false>);
  case 4:
    return format == 1 ? f(&k_batch_step<4, true>) : f(&k_batch_step<4, false>);
  case 8:
    return format == 1 ? f(&k_batch_step<8, true>) : f(&k_batch_step<8, false>);
  default:
    return TRL_HIP_EINVAL;
  }
}

bool valid_rpt(int rpt) { return rpt == 1 || rpt == 2 || rpt == 4 || rpt == 8 || rpt == 16 || rpt == 32; }

// Builds the tiles for rows-per-thread rpt and checks that the whole grid
// fits on the device at once (the barriers need every tile resident).
int build_tiles(trl_hip_batch* bt, int rpt, bool all = false)
{
  const int rows = kT * rpt;
  std::vector<int> tp, r0, r1, loc, pn(bt->nprob), pt0(bt->nprob);
  for (int p = 0; p < bt->nprob; ++p) {
    pt0[p] = int(tp.size());
    const size_t np = bt->n[p];
    // active problems only (a step or rotation never touches the others);
    // every problem for the residual pass
    const int cnt = (all || bt->active_h[p]) ? int((np + rows - 1) / rows) : 0;
    pn[p] = cnt;
    for (int k = 0; k < cnt; ++k) {
      tp.push_back(p);
      r0.push_back(int(bt->roff[p] + size_t(k) * rows));
      r1.push_back(int(std::min(bt->roff[p] + size_t(k + 1) * rows, bt->roff[p] + np)));
      loc.push_back(k);
    }
  }
  const int nt = int(tp.size());
  int occ = 0;
  const size_t shm = size_t(bt->m + 1) * bt->b * bt->b * sizeof(double);
  int rc = with_step_kernel(bt->b, bt->format, [&](auto kfn) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, kT, shm) == hipSuccess ? 0 : TRL_HIP_EINVAL;
  });
  if (rc) return rc;
  // the step kernel's barriers need every tile resident at once
  if (!all && (occ < 1 || size_t(nt) > size_t(occ) * size_t(bt->ctx->num_cus))) return TRL_HIP_ERANGE;
  dfree(bt->tile_prob);
  dfree(bt->tile_row0);
  dfree(bt->tile_row1);
  dfree(bt->tile_local);
  dfree(bt->prob_ntiles);
  dfree(bt->prob_tile0);
  if ((rc = dalloc(&bt->tile_prob, nt)) || (rc = dalloc(&bt->tile_row0, nt)) || (rc = dalloc(&bt->tile_row1, nt)) || (rc = dalloc(&bt->tile_local, nt)) ||
      (rc = dalloc(&bt->prob_ntiles, bt->nprob)) || (rc = dalloc(&bt->prob_tile0, bt->nprob)))
    return rc;
  TRL_HIP_TRY(hipMemcpy(bt->tile_prob, tp.data(), nt * sizeof(int), hipMemcpyHostToDevice));
  TRL_HIP_TRY(hipMemcpy(bt->tile_row0, r0.data(), nt * sizeof(int), hipMemcpyHostToDevice));
  TRL_HIP_TRY(hipMemcpy(bt->tile_row1, r1.data(), nt * sizeof(int), hipMemcpyHostToDevice));
  TRL_HIP_TRY(hipMemcpy(bt->tile_local, loc.data(), nt * sizeof(int), hipMemcpyHostToDevice));
  TRL_HIP_TRY(hipMemcpy(bt->prob_ntiles, pn.data(), bt->nprob * sizeof(int), hipMemcpyHostToDevice));
  TRL_HIP_TRY(hipMemcpy(bt->prob_tile0, pt0.data(), bt->nprob * sizeof(int), hipMemcpyHostToDevice));
  const size_t need = size_t(2) * nt * (bt->m + 1) * bt->b * bt->b;
  if (need > bt->partials_cap) {
    dfree(bt->partials);
    if ((rc = dalloc(&bt->partials, need))) return rc;
    bt->partials_cap = need;
  }
  bt->ntiles = nt;
  bt->rpt = rpt;
  bt->tile_prob_h = tp;
  bt->tile_row0_h = r0;
  bt->tile_row1_h = r1;
  bt->prob_ntiles_h = pn;
  bt->prob_tile0_h = pt0;
  clear_graphs(bt);
  return 0;
}

// Enqueue the launches of steps k..m_end-1 (k == 0: orthonormalise block 0
// first; quasi: the first step is the quasi-Lanczos step).
int enqueue_cycle(trl_hip_batch* bt, int k, int m_end, int quasi)
{
  if (bt->ntiles == 0) return 0;
  const BatchDev d = make_dev(bt);
  const size_t nc = size_t(bt->nprob) * (bt->m + 1);
  hipLaunchKernelGGL(k_zero_u32, dim3(unsigned(std::min<size_t>((nc + 255) / 256, 64))), dim3(256), 0, bt->ctx->stream, bt->counters, nc);
  TRL_HIP_LAUNCH_CHECK();
  const size_t shm = size_t(bt->m + 1) * bt->b * bt->b * sizeof(double);
  int seq = 0;
  return with_step_kernel(bt->b, bt->format, [&](auto kfn) -> int {
    auto launch = [&](int i, int kind) -> int {
      hipLaunchKernelGGL(kfn, dim3(unsigned(bt->ntiles)), dim3(kT), shm, bt->ctx->stream, d, i, kind, seq, bt->cq_passes, bt->adaptive);
      ++seq;
      TRL_HIP_LAUNCH_CHECK();
      return 0;
    };
    int rc = 0;
    if (k == 0 && (rc = launch(0, 2))) return rc;
    for (int i = k; i < m_end; ++i)
      if ((rc = launch(i, (quasi && i == k) ? 1 : 0))) return rc;
    return 0;
  });
}

} // namespace

extern "C" {

int trl_hip_batch_create(trl_hip_ctx* ctx, int nprob, const size_t* n, const int32_t* const* rowptr, const int32_t* const* col, const double* const* val,
                         int b, int ncv, trl_hip_batch** out)
{
  if (!ctx || !out || nprob < 1 || !n || !rowptr || !col || !val || b < 1 || b > kMaxB || !valid_b(b) || b == 6 || ncv < b || ncv % b) return TRL_HIP_EINVAL;
  *out = nullptr;
  auto* bt = new trl_hip_batch();
  bt->ctx = ctx;
  bt->nprob = nprob;
  bt->b = b;
  bt->m = ncv / b;
  bt->n.assign(n, n + nprob);
  bt->roff.resize(nprob);
  size_t N = 0;
  long long nnz = 0;
  int width = 0;
  for (int p = 0; p < nprob; ++p) {
    if (!rowptr[p] || (n[p] && (!col[p] || !val[p])) || n[p] == 0) {
      delete bt;
      return TRL_HIP_EINVAL;
    }
    bt->roff[p] = N;
    N += n[p];
    nnz += rowptr[p][n[p]];
    for (size_t r = 0; r < n[p]; ++r) width = std::max(width, int(rowptr[p][r + 1] - rowptr[p][r]));
  }
  if (N > size_t(INT32_MAX) || nnz > INT32_MAX) {
    delete bt;
    return TRL_HIP_EINVAL;
  }
  bt->N = N;
  bt->nnz = nnz;
  bt->width = width;
  std::vector<int32_t> rp(N + 1), ci(nnz);
  std::vector<double> vv(nnz);
  size_t row = 0;
  long long pos = 0;
  for (int p = 0; p < nprob; ++p) {
    for (size_t r = 0; r < n[p]; ++r, ++row) {
      rp[row] = int32_t(pos);
      for (int32_t q = rowptr[p][r]; q < rowptr[p][r + 1]; ++q, ++pos) {
        ci[pos] = int32_t(bt->roff[p] + size_t(col[p][q]));
        vv[pos] = val[p][q];
      }
    }
  }
  rp[N] = int32_t(pos);
  int rc = 0;
  auto fail = [&](int code) {
    trl_hip_batch_destroy(bt);
    return code;
  };
  if ((rc = dalloc(&bt->rowptr, N + 1)) || (rc = dalloc(&bt->col, nnz)) || (rc = dalloc(&bt->val, nnz))) return fail(rc);
  if (hipMemcpy(bt->rowptr, rp.data(), (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(bt->col, ci.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(bt->val, vv.data(), nnz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(TRL_HIP_EINVAL);
  if ((rc = dalloc(&bt->ell_col, size_t(width) * N)) || (rc = dalloc(&bt->ell_val, size_t(width) * N))) return fail(rc);
  if ((rc = trl_hip_csr_to_ell(ctx, TRL_HIP_F64, N, bt->rowptr, bt->col, bt->val, width, bt->ell_col, bt->ell_val))) return fail(rc);
  const size_t basis = size_t(bt->m + 1) * N * b;
  const int BB = b * b;
  if ((rc = dalloc(&bt->V[0], basis)) || (rc = dalloc(&bt->V[1], basis)) || (rc = dalloc(&bt->vsel, 1)) || (rc = dalloc(&bt->active, nprob)) ||
      (rc = dalloc(&bt->counters, size_t(nprob) * (bt->m + 1))) || (rc = dalloc(&bt->alpha, size_t(nprob) * bt->m * BB)) ||
      (rc = dalloc(&bt->beta, size_t(nprob) * bt->m * BB)) || (rc = dalloc(&bt->beta_prev, size_t(nprob) * BB)) ||
      (rc = dalloc(&bt->arrow, size_t(nprob) * BB)) || (rc = dalloc(&bt->pflag, nprob)) || (rc = dalloc(&bt->err, 1)))
    return fail(rc);
  if (hipMemset(bt->V[0], 0, basis * sizeof(double)) != hipSuccess || hipMemset(bt->V[1], 0, basis * sizeof(double)) != hipSuccess ||
      hipMemset(bt->vsel, 0, sizeof(int)) != hipSuccess || hipMemset(bt->pflag, 0, nprob * sizeof(int)) != hipSuccess ||
      hipMemset(bt->err, 0, sizeof(int)) != hipSuccess || hipMemset(bt->arrow, 0, size_t(nprob) * BB * sizeof(double)) != hipSuccess ||
      hipMemset(bt->beta_prev, 0, size_t(nprob) * BB * sizeof(double)) != hipSuccess)
    return fail(TRL_HIP_EINVAL);
  bt->active_h.assign(nprob, 1);
  bt->saved_h.assign(nprob, 0);
  if (hipMemcpy(bt->active, bt->active_h.data(), nprob * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return fail(TRL_HIP_EINVAL);
  // default variant: CSR, the largest rows-per-thread whose grid fits
  bt->format = 0;
  int ok = -1;
  for (int rpt : {4, 8, 16, 32})
    if (build_tiles(bt, rpt) == 0) {
      ok = rpt;
      break;
    }
  if (ok < 0) return fail(TRL_HIP_ERANGE);
  if (hipDeviceSynchronize() != hipSuccess) return fail(TRL_HIP_EINVAL);
  *out = bt;
  return 0;
}

int trl_hip_batch_destroy(trl_hip_batch* bt)
{
  if (!bt) return TRL_HIP_EINVAL;
  (void)hipStreamSynchronize(bt->ctx->stream);
  clear_graphs(bt);
  dfree(bt->rowptr);
  dfree(bt->col);
  dfree(bt->val);
  dfree(bt->ell_col);
  dfree(bt->ell_val);
  dfree(bt->tile_prob);
  dfree(bt->tile_row0);
  dfree(bt->tile_row1);
  dfree(bt->tile_local);
  dfree(bt->prob_ntiles);
  dfree(bt->prob_tile0);
  dfree(bt->V[0]);
  dfree(bt->V[1]);
  dfree(bt->vsel);
  dfree(bt->active);
  dfree(bt->partials);
  dfree(bt->counters);
  dfree(bt->alpha);
  dfree(bt->beta);
  dfree(bt->beta_prev);
  dfree(bt->arrow);
  dfree(bt->Ybuf);
  dfree(bt->lam);
  dfree(bt->rsum);
  dfree(bt->Xres);
  dfree(bt->pflag);
  dfree(bt->err);
  delete bt;
  return 0;
}

int trl_hip_batch_set_variant(trl_hip_batch* bt, int format, int rows_per_thread)
{
  if (!bt || (format != 0 && format != 1) || !valid_rpt(rows_per_thread)) return TRL_HIP_EINVAL;
  const int old_format = bt->format, old_rpt = bt->rpt;
  bt->format = format;
  const int rc = build_tiles(bt, rows_per_thread);
  if (rc) {
    bt->format = old_format;
    (void)build_tiles(bt, old_rpt);
  }
  return rc;
}

int trl_hip_batch_info(const trl_hip_batch* bt, int* ntiles, int* format, int* rows_per_thread, double* tuned_us)
{
  if (!bt) return TRL_HIP_EINVAL;
  if (ntiles) *ntiles = bt->ntiles;
  if (format) *format = bt->format;
  if (rows_per_thread) *rows_per_thread = bt->rpt;
  if (tuned_us) *tuned_us = bt->tuned_us;
  return 0;
}

int trl_hip_batch_set_options(trl_hip_batch* bt, int cholqr_passes, int use_graph)
{
  if (!bt || cholqr_passes < 1 || cholqr_passes > 2) return TRL_HIP_EINVAL;
  if (bt->cq_passes != cholqr_passes) clear_graphs(bt);
  bt->cq_passes = cholqr_passes;
  bt->use_graph = use_graph != 0;
  return 0;
}

int trl_hip_batch_set_reorth(trl_hip_batch* bt, int adaptive)
{
  if (!bt || (adaptive != 0 && adaptive != 1)) return TRL_HIP_EINVAL;
  if (bt->adaptive != adaptive) clear_graphs(bt);
  bt->adaptive = adaptive;
  return 0;
}

int trl_hip_batch_set_block(trl_hip_batch* bt, int p, int q, const double* host)
{
  if (!bt || p < 0 || p >= bt->nprob || q < 0 || q > bt->m || !host) return TRL_HIP_EINVAL;
  double* dst = bt->V[bt->vsel_h] + (size_t(q) * bt->N + bt->roff[p]) * bt->b;
  TRL_HIP_TRY(hipMemcpyAsync(dst, host, bt->n[p] * bt->b * sizeof(double), hipMemcpyHostToDevice, bt->ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(bt->ctx->stream));
  return 0;
}

int trl_hip_batch_get_block(trl_hip_batch* bt, int p, int q, double* host)
{
  if (!bt || p < 0 || p >= bt->nprob || q < 0 || q > bt->m || !host) return TRL_HIP_EINVAL;
  const bool saved = bt->saved_h[p] && q < bt->nblk_saved;
  const double* base = saved ? bt->Xres : bt->V[bt->vsel_h];
  const double* src = base + (size_t(q) * bt->N + bt->roff[p]) * bt->b;
  TRL_HIP_TRY(hipMemcpyAsync(host, src, bt->n[p] * bt->b * sizeof(double), hipMemcpyDeviceToHost, bt->ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(bt->ctx->stream));
  return 0;
}

int trl_hip_batch_set_active(trl_hip_batch* bt, const int* active)
{
  if (!bt || !active) return TRL_HIP_EINVAL;
  bool need_tiles = false;
  for (int p = 0; p < bt->nprob; ++p) {
    bt->active_h[p] = active[p] ? 1 : 0;
    if (bt->active_h[p] && bt->prob_ntiles_h[p] == 0) need_tiles = true;
  }
  TRL_HIP_TRY(hipMemcpyAsync(bt->active, bt->active_h.data(), bt->nprob * sizeof(int), hipMemcpyHostToDevice, bt->ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(bt->ctx->stream));
  // a problem that (re)joins needs tiles; finished ones keep theirs until the
  // next retiling (trl_hip_batch_set_variant / trl_hip_batch_tune)
  return need_tiles ? build_tiles(bt, bt->rpt) : 0;
}

int trl_hip_batch_extend(trl_hip_batch* bt, int k, int m_end, int quasi, const double* arrow, double* alpha, double* beta, int* pflag)
{
  if (!bt || k < 0 || m_end > bt->m || k >= m_end || (quasi && (k == 0 || !arrow))) return TRL_HIP_EINVAL;
  const int BB = bt->b * bt->b;
  hipStream_t s = bt->ctx->stream;
  if (quasi) TRL_HIP_TRY(hipMemcpyAsync(bt->arrow, arrow, size_t(bt->nprob) * BB * sizeof(double), hipMemcpyHostToDevice, s));
  int rc = 0;
  if (bt->use_graph) {
    const auto key = std::make_tuple(k, m_end, quasi);
    auto it = bt->graphs.find(key);
    if (it == bt->graphs.end()) {
      hipGraph_t g = nullptr;
      TRL_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      rc = enqueue_cycle(bt, k, m_end, quasi);
      const hipError_t e = hipStreamEndCapture(s, &g);
      if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : int(e);
      }
      hipGraphExec_t ge = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) return int(ei);
      it = bt->graphs.emplace(key, ge).first;
    }
    TRL_HIP_TRY(hipGraphLaunch(it->second, s));
  }
  else if ((rc = enqueue_cycle(bt, k, m_end, quasi)))
    return rc;
  // the copies land in host memory owned by this frame and the caller:
  // always wait for the stream before returning, also on an error
  int err = 0;
  hipError_t e = hipMemcpyAsync(&err, bt->err, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && alpha) e = hipMemcpyAsync(alpha, bt->alpha, size_t(bt->nprob) * bt->m * BB * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && beta) e = hipMemcpyAsync(beta, bt->beta, size_t(bt->nprob) * bt->m * BB * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && pflag) e = hipMemcpyAsync(pflag, bt->pflag, bt->nprob * sizeof(int), hipMemcpyDeviceToHost, s);
  const hipError_t es = hipStreamSynchronize(s);
  if (e != hipSuccess) return int(e);
  if (es != hipSuccess) return int(es);
  return err ? TRL_HIP_EBARRIER : 0;
}

int trl_hip_batch_rotate(trl_hip_batch* bt, const double* Y, int kr)
{
  if (!bt || !Y || kr < 1 || kr >= bt->m) return TRL_HIP_EINVAL;
  if (bt->ntiles == 0) return 0;
  const int b = bt->b;
  const size_t ysz = size_t(bt->m) * b * kr * b;
  const size_t shm = ysz * sizeof(double);
  if (shm > bt->ctx->max_lds) return TRL_HIP_ERANGE;
  hipStream_t s = bt->ctx->stream;
  // the coefficient buffer keeps its size (kr is the same every cycle): no
  // allocation, and no free (which would wait for the device), per cycle
  if (bt->ybuf_count < ysz * bt->nprob) {
    dfree(bt->Ybuf);
    bt->ybuf_count = 0;
    const int rc = dalloc(&bt->Ybuf, ysz * bt->nprob);
    if (rc) return rc;
    bt->ybuf_count = ysz * bt->nprob;
  }
  TRL_HIP_TRY(hipMemcpyAsync(bt->Ybuf, Y, ysz * bt->nprob * sizeof(double), hipMemcpyHostToDevice, s));
  const BatchDev d = make_dev(bt);
  switch (b) {
  case 1:
    hipLaunchKernelGGL((k_batch_rotate<1>), dim3(unsigned(bt->ntiles)), dim3(kT), shm, s, d, bt->Ybuf, kr);
    break;
  case 2:
    hipLaunchKernelGGL((k_batch_rotate<2>), dim3(unsigned(bt->ntiles)), dim3(kT), shm, s, d, bt->Ybuf, kr);
    break;
  case 4:
    hipLaunchKernelGGL((k_batch_rotate<4>), dim3(unsigned(bt->ntiles)), dim3(kT), shm, s, d, bt->Ybuf, kr);
    break;
  case 8:
    hipLaunchKernelGGL((k_batch_rotate<8>), dim3(unsigned(bt->ntiles)), dim3(kT), shm, s, d, bt->Ybuf, kr);
    break;
  default:
    return TRL_HIP_EINVAL;
  }
  TRL_HIP_LAUNCH_CHECK();
  bt->vsel_h ^= 1;
  TRL_HIP_TRY(hipMemcpyAsync(bt->vsel, &bt->vsel_h, sizeof(int), hipMemcpyHostToDevice, s));
  TRL_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int trl_hip_batch_save_blocks(trl_hip_batch* bt, int p, int nblk)
{
  if (!bt || p < 0 || p >= bt->nprob || nblk < 1 || nblk > bt->m + 1) return TRL_HIP_EINVAL;
  int rc = 0;
  if (!bt->Xres) {
    if ((rc = dalloc(&bt->Xres, size_t(nblk) * bt->N * bt->b))) return rc;
    TRL_HIP_TRY(hipMemsetAsync(bt->Xres, 0, size_t(nblk) * bt->N * bt->b * sizeof(double), bt->ctx->stream));
    bt->nblk_saved = nblk;
  }
  if (nblk > bt->nblk_saved) return TRL_HIP_EINVAL;
  const double* V = bt->V[bt->vsel_h];
  for (int q = 0; q < nblk; ++q) {
    const size_t off = (size_t(q) * bt->N + bt->roff[p]) * bt->b;
    TRL_HIP_TRY(hipMemcpyAsync(bt->Xres + off, V + off, bt->n[p] * bt->b * sizeof(double), hipMemcpyDeviceToDevice, bt->ctx->stream));
  }
  TRL_HIP_TRY(hipStreamSynchronize(bt->ctx->stream));
  bt->saved_h[p] = 1;
  return 0;
}

int trl_hip_batch_residuals(trl_hip_batch* bt, int nblk, const double* lambda, double* out)
{
  if (!bt || nblk < 1 || nblk > bt->m + 1 || !lambda || !out) return TRL_HIP_EINVAL;
  int rc = 0;
  // every problem's leading blocks in the saved array (those still in the basis now)
  for (int p = 0; p < bt->nprob; ++p)
    if (!bt->saved_h[p] && (rc = trl_hip_batch_save_blocks(bt, p, nblk))) return rc;
  if (nblk > bt->nblk_saved) return TRL_HIP_EINVAL;
  // tiles over every problem for this pass (no barriers: no residency limit)
  const int step_rpt = bt->rpt;
  if ((rc = build_tiles(bt, step_rpt, true))) return rc;
  const int b = bt->b;
  hipStream_t s = bt->ctx->stream;
  const size_t nl = size_t(bt->nprob) * nblk * b;
  const size_t ns = size_t(bt->ntiles) * nblk * 2 * b;
  dfree(bt->lam);
  dfree(bt->rsum);
  if ((rc = dalloc(&bt->lam, nl)) || (rc = dalloc(&bt->rsum, ns))) return rc;
  TRL_HIP_TRY(hipMemcpyAsync(bt->lam, lambda, nl * sizeof(double), hipMemcpyHostToDevice, s));
  const BatchDev d = make_dev(bt);
  const bool ell = bt->format == 1;
#define TRL_BRES(B_)                                                                                                                                        \
  if (ell)                                                                                                                                                  \
    hipLaunchKernelGGL((k_batch_residual<B_, true>), dim3(unsigned(bt->ntiles)), dim3(kT), 0, s, d, bt->Xres, nblk, bt->lam, bt->rsum);                  \
  else                                                                                                                                                      \
    hipLaunchKernelGGL((k_batch_residual<B_, false>), dim3(unsigned(bt->ntiles)), dim3(kT), 0, s, d, bt->Xres, nblk, bt->lam, bt->rsum);
  switch (b) {
  case 1:
    TRL_BRES(1);
    break;
  case 2:
    TRL_BRES(2);
    break;
  case 4:
    TRL_BRES(4);
    break;
  case 8:
    TRL_BRES(8);
    break;
  default:
    return TRL_HIP_EINVAL;
  }
#undef TRL_BRES
  TRL_HIP_LAUNCH_CHECK();
  std::vector<double> sums(ns);
  const hipError_t e = hipMemcpyAsync(sums.data(), bt->rsum, ns * sizeof(double), hipMemcpyDeviceToHost, s);
  const hipError_t es = hipStreamSynchronize(s);
  if (e != hipSuccess) return int(e);
  if (es != hipSuccess) return int(es);
  std::vector<double> acc(size_t(bt->nprob) * nblk * 2 * b, 0.0);
  for (int t = 0; t < bt->ntiles; ++t) {
    const int p = bt->tile_prob_h[t];
    for (size_t e = 0; e < size_t(nblk) * 2 * b; ++e) acc[size_t(p) * nblk * 2 * b + e] += sums[size_t(t) * nblk * 2 * b + e];
  }
  for (int p = 0; p < bt->nprob; ++p)
    for (int q = 0; q < nblk; ++q)
      for (int c = 0; c < b; ++c) {
        const double* a = acc.data() + (size_t(p) * nblk + q) * 2 * b;
        out[(size_t(p) * nblk + q) * b + c] = std::sqrt(a[c]) / std::max(std::sqrt(a[b + c]), 1e-300);
      }
  (void)build_tiles(bt, step_rpt); // back to the active problems' tiles
  return 0;
}

int trl_hip_batch_raise_error(trl_hip_batch* bt)
{
  if (!bt) return TRL_HIP_EINVAL;
  const int one = 1;
  TRL_HIP_TRY(hipMemcpyAsync(bt->err, &one, sizeof(int), hipMemcpyHostToDevice, bt->ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(bt->ctx->stream));
  return 0;
}

int trl_hip_batch_reset_flags(trl_hip_batch* bt)
{
  if (!bt) return TRL_HIP_EINVAL;
  TRL_HIP_TRY(hipMemsetAsync(bt->pflag, 0, bt->nprob * sizeof(int), bt->ctx->stream));
  TRL_HIP_TRY(hipMemsetAsync(bt->err, 0, sizeof(int), bt->ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(bt->ctx->stream));
  return 0;
}

int trl_hip_batch_tune(trl_hip_batch* bt, double* best_us)
{
  if (!bt) return TRL_HIP_EINVAL;
  // time one full-width step (the last step of a cycle: reorthogonalisation
  // against m blocks) of every variant whose grid fits, on the current basis
  // contents; the basis and the coefficient arrays are scratch here
  hipStream_t s = bt->ctx->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return TRL_HIP_ENOMEM;
  const int old_format = bt->format, old_rpt = bt->rpt;
  const bool old_graph = bt->use_graph;
  bt->use_graph = false;
  double best = 1e300;
  int bf = old_format, br = old_rpt;
  for (int format : {0, 1})
    for (int rpt : {1, 2, 4, 8, 16}) {
      bt->format = format;
      if (build_tiles(bt, rpt)) continue;
      float ms = 0.f;
      bool ok = true;
      for (int rep = 0; rep < 3 && ok; ++rep) {
        ok = hipEventRecord(e0, s) == hipSuccess && enqueue_cycle(bt, bt->m - 1, bt->m, 0) == 0 && hipEventRecord(e1, s) == hipSuccess &&
             hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      }
      int err = 0;
      ok = ok && hipMemcpy(&err, bt->err, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess && err == 0;
      if (ok && ms * 1e3 < best) {
        best = ms * 1e3;
        bf = format;
        br = rpt;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  bt->use_graph = old_graph;
  bt->format = bf;
  int rc = build_tiles(bt, br);
  if (!rc) rc = trl_hip_batch_reset_flags(bt);
  bt->tuned_us = best < 1e300 ? best : 0.0;
  if (best_us) *best_us = bt->tuned_us;
  return rc;
}

} // extern "C"
