// Sparse operator kernels: Y = diag(d) A X (or the fused three-term form
// Y = alpha diag(d) A X + beta X + gamma Z) for n x b row-major panels.
//
// Row kernels (one row per CS lanes, CS in {1, 2, 4}; lane `part` owns
// columns [part*b/CS, (part+1)*b/CS) of the row): the row's column indices and
// values are loaded kChunk at a time before the gathers of X, so a chunk costs
// two memory latencies instead of two per nonzero; padding slots gather the
// row's own X row with value 0.  Storage is CSR (int32 rowptr/col) or
// slot-major ELL (slot s of row r at s*n + r; fully coalesced index/value
// streams).  Splitting columns instead of nonzeros keeps every lane's
// registers small (b/CS values per gathered row) without a reduction.
//
// Lane-group kernel (CSR only, L in {2,4,8,16} lanes per row): lanes walk the
// row's nonzeros sub, sub+L, ... and the partial rows are combined by a
// butterfly (shfl_xor) inside the group; for long rows.
//
// Several panels in one launch: every lane applies its row to each panel in
// turn; a single-chunk ELL row keeps its indices and values in registers for
// all of them, so the matrix is read once per launch, not once per panel.
//
// Workgroups are renumbered so that each XCD walks a contiguous range of rows
// (in index order, or plane block by plane block, trl_hip_spmat::plane): the
// stencil neighbours of a row (r +- 1, r +- nx, ...) then share that XCD's L2.
// The ELL matrix stream (indices, values; every cache line used by one load),
// the row scaling, Z and Y are touched once and use non-temporal accesses,
// which leaves the L2 to the gathered rows of X.  CSR index and value lines
// are shared by consecutive loads of neighbouring lanes and stay cached.
#include <algorithm>
#include <type_traits>

#include "common.hh"

using namespace trl_hip_dev;

namespace {

constexpr int kThreads = 256;
// Workgroup size of the product kernels.  FP32 lanes holding at most 4
// values of a row need under 64 registers, so two 1024-lane workgroups fit a
// CU; so few workgroup launches (millions of rows, one or a few lanes each)
// keep dispatch from limiting the memory stream (measured: 85 us against
// 95 us with 256 lanes for the FP32 ELL product of a 4M-row 7-point stencil,
// b = 8).  Wider lanes take 256: with one 1024-lane workgroup per CU the
// whole CU waits for its slowest wave.
template <typename S, int BC>
constexpr int spmm_threads()
{
  return (sizeof(S) == 4 && BC <= 4) ? 1024 : 256;
}
constexpr unsigned kXcds = 8;
constexpr int kChunk = 8;
// forcing coefficients one launch of the several-panel kernel stages in LDS
constexpr int kForceCoefs = 256;

__device__ inline unsigned xcd_block(unsigned bid, unsigned nb)
{
  if (nb % kXcds != 0) return bid;
  return (bid % kXcds) * (nb / kXcds) + bid / kXcds;
}

struct Epi {
  const void* dscale;
  const void* Z;
  double alpha, beta, gamma;
  unsigned npan;  // panels of X, Y, Z processed by one launch
  size_t pstride; // elements between consecutive panels
  // row order: plane = 0 index order; else blocks of pblock rows of each of
  // the nz planes in turn (trl_hip_spmat::plane / plane_block)
  unsigned plane, pblock, nz;
  // forcing term of the axpby form (F may be null): + F_p diag(g_p), F laid
  // out as X, g[p * B + c] the coefficient of column c of panel p
  const void* F;
  const double* g;
  // tiled order inside the plane blocks (tly > 0): a plane holds lines of
  // `line` rows, and a workgroup takes tly consecutive lines of
  // rows_per_wg / tly rows each (the launcher checks that the tiles cover
  // every plane block exactly)
  unsigned line = 0, tly = 0;
  // several-panel ELL step: in-wave gathers from the owning lane
  unsigned share = 0;
};

// Row handled by local row slot `slot` (0 .. rows_per_wg-1) of row-range
// workgroup wrow, or -1 when that slot has no row (past n, or past the end
// of its plane block).  Blocked order: a segment is one block of one plane,
// covered by ceil(pblock / rows_per_wg) consecutive workgroups; segments run
// through the planes of block 0, then of block 1, ...
__device__ inline int row_of(unsigned wrow, unsigned slot, unsigned rows_per_wg, int n, const Epi& e)
{
  if (e.plane == 0) {
    const unsigned r = wrow * rows_per_wg + slot;
    return r < unsigned(n) ? int(r) : -1;
  }
  const unsigned wps = (e.pblock + rows_per_wg - 1) / rows_per_wg;
  if (e.tly) {
    const unsigned tx = rows_per_wg / e.tly, xt = e.line / tx;
    const unsigned seg = wrow / wps, w_in = wrow % wps;
    const unsigned pb = seg / e.nz, z = seg % e.nz;
    const unsigned y = (w_in / xt) * e.tly + slot / tx, x = (w_in % xt) * tx + slot % tx;
    const unsigned p = pb * e.pblock + y * e.line + x;
    if (p >= e.plane) return -1;
    const size_t r = size_t(z) * e.plane + p;
    return r < size_t(n) ? int(r) : -1;
  }
  const unsigned seg = wrow / wps, in_seg = (wrow % wps) * rows_per_wg + slot;
  const unsigned pb = seg / e.nz, z = seg % e.nz;
  const unsigned p = pb * e.pblock + in_seg;
  if (in_seg >= e.pblock || p >= e.plane) return -1;
  const size_t r = size_t(z) * e.plane + p;
  return r < size_t(n) ? int(r) : -1;
}

// Address of element i of an array.  O32: i * sizeof(T) fits in 32 bits
// (checked by the launcher), so the address is a 64-bit workgroup-uniform
// base plus a 32-bit lane offset, which the global loads take as an SGPR
// base and a VGPR offset: one VGPR per address instead of two, and a whole
// ELL chunk of slot loads shares one offset register.
template <bool O32>
using Idx = std::conditional_t<O32, unsigned, size_t>;

template <bool O32, typename T>
__device__ inline T* elem(T* base, Idx<O32> i)
{
  if constexpr (O32) {
    using C = std::conditional_t<std::is_const_v<T>, const char, char>;
    return reinterpret_cast<T*>(reinterpret_cast<C*>(base) + i * unsigned(sizeof(T)));
  }
  else
    return base + i;
}

// AXPBY = false: Y = d*acc; true: Y = alpha*d*acc + beta*X + gamma*Z
// + F diag(g) (Z, F may be null).  X, Y, Z, F point at the lane's panel;
// ro = row * B + off, the first element of this lane in it; g (with F) at
// the coefficient of the lane's first column.
template <typename S, int BC, bool AXPBY, bool O32>
__device__ inline void epilogue(const S (&acc)[BC], Idx<O32> row, Idx<O32> ro, const S* __restrict__ X, S* Y, const S* Z, const S* F, const double* g,
                                const Epi& e)
{
  const S* ds = static_cast<const S*>(e.dscale);
  S out[BC];
  if constexpr (!AXPBY) {
    const S d = ds ? ld<true>(elem<O32>(ds, row)) : S(1);
#pragma unroll
    for (int q = 0; q < BC; ++q) out[q] = ds ? acc[q] * d : acc[q];
  }
  else {
    S xr[BC], zr[BC], fr[BC];
    RowIO<S, BC>::load(elem<O32>(X, ro), xr);
    if (Z) RowIO<S, BC>::template load<true>(elem<O32>(Z, ro), zr);
    if (F) RowIO<S, BC>::template load<true>(elem<O32>(F, ro), fr);
    const double a = ds ? e.alpha * double(ld<true>(elem<O32>(ds, row))) : e.alpha;
#pragma unroll
    for (int q = 0; q < BC; ++q) {
      double v = a * double(acc[q]) + e.beta * double(xr[q]);
      if (Z) v += e.gamma * double(zr[q]);
      if (F) v += g[q] * double(fr[q]);
      out[q] = S(v);
    }
  }
  RowIO<S, BC>::template store<true>(elem<O32>(Y, ro), out);
}

template <typename S, int B, int CS, bool ELL, bool AXPBY, bool O32>
__global__ __launch_bounds__((spmm_threads<S, B / CS>())) void k_spmm_rows(const int32_t* __restrict__ rowptr, int width, const int32_t* __restrict__ col,
                                                                           const S* __restrict__ val, const S* __restrict__ X, S* Y, Epi e, int n)
{
  using I = Idx<O32>;
  constexpr int BC = B / CS;
  const int row = row_of(xcd_block(blockIdx.x, gridDim.x), threadIdx.x / CS, blockDim.x / CS, n, e);
  const int off = int(threadIdx.x % CS) * BC;
  if (row < 0) return;
  S acc[BC];
#pragma unroll
  for (int c = 0; c < BC; ++c) acc[c] = S(0);
  if constexpr (ELL) {
    // slot-major ELL: slot s of every row at s*n + row.  The slot count is
    // uniform, so padding slots of the last chunk are skipped by every lane.
    for (int s0 = 0; s0 < width; s0 += kChunk) {
      int ci[kChunk];
      S vi[kChunk];
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
        if (s0 + q < width) {
          const size_t base = size_t(s0 + q) * size_t(n);
          ci[q] = ld<true>(elem<O32>(col + base, I(row))); // every line used by one load: stream it
          vi[q] = ld<true>(elem<O32>(val + base, I(row)));
        }
      S x[kChunk][BC];
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
        if (s0 + q < width) RowIO<S, BC>::load(elem<O32>(X, I(ci[q]) * I(B) + I(off)), x[q]);
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
        if (s0 + q < width)
#pragma unroll
          for (int c = 0; c < BC; ++c) acc[c] = fma(vi[q], x[q][c], acc[c]);
    }
  }
  else {
    const int beg = rowptr[row];
    const int end = rowptr[row + 1];
    for (int p0 = beg; p0 < end; p0 += kChunk) {
      int ci[kChunk];
      S vi[kChunk];
#pragma unroll
      for (int q = 0; q < kChunk; ++q) {
        const bool ok = p0 + q < end;
        ci[q] = ok ? col[p0 + q] : row; // padding gathers the row's own X row with value 0
        vi[q] = ok ? val[p0 + q] : S(0);
      }
      S x[kChunk][BC];
#pragma unroll
      for (int q = 0; q < kChunk; ++q) RowIO<S, BC>::load(elem<O32>(X, I(ci[q]) * I(B) + I(off)), x[q]);
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
#pragma unroll
        for (int c = 0; c < BC; ++c) acc[c] = fma(vi[q], x[q][c], acc[c]);
    }
  }
  epilogue<S, BC, AXPBY, O32>(acc, I(row), I(row) * I(B) + I(off), X, Y, static_cast<const S*>(e.Z), static_cast<const S*>(e.F), e.g ? e.g + off : nullptr,
                              e);
}

// Several panels of an ELL matrix of at most kChunk slots: each lane loads
// its row's indices and values once and applies them to every panel in turn
// (panel p of X, Y, Z at p * pstride elements), so the matrix is read from
// HBM once per launch instead of once per panel.  Working through the panels
// one after another keeps one panel's X, Z and Y streams in flight at a time
// (measured faster than spreading a workgroup's lanes over all panels at
// once: nine concurrent write streams cost more than the matrix saves).
// Workgroup size of the several-panel kernel: 256 lanes (measured on the
// 4M-row stencil, b = 8, nine panels: 1024-lane workgroups take 8% longer).
template <typename S, int BC>
constexpr int panel_threads()
{
  return 256;
}

template <typename S, int B, int CS, bool AXPBY, bool O32, bool FORCE>
__global__ __launch_bounds__((panel_threads<S, B / CS>())) void k_spmm_ell_panels(int width, const int32_t* __restrict__ col, const S* __restrict__ val, const S* __restrict__ X, S* Y,
                                                         Epi e, int n)
{
  using I = Idx<O32>;
  constexpr int BC = B / CS;
  // the forcing coefficients of all panels, staged in LDS (npan * B <=
  // kForceCoefs, checked by the launcher): read just before their use, they
  // hold no registers across the gathers
  __shared__ double gsh[FORCE ? kForceCoefs : 1];
  if constexpr (FORCE) {
    for (unsigned i = threadIdx.x; i < e.npan * B; i += blockDim.x) gsh[i] = e.g[i];
    __syncthreads();
  }
  const int row = row_of(xcd_block(blockIdx.x, gridDim.x), threadIdx.x / CS, blockDim.x / CS, n, e);
  const int off = int(threadIdx.x % CS) * BC;
  if (row < 0) return;
  const I ro = I(row) * I(B) + I(off);
  const S* Z = static_cast<const S*>(e.Z);
  // the forcing term has an instantiation of its own (its registers would
  // cost the plain step a wave per SIMD)
  const S* F = FORCE ? static_cast<const S*>(e.F) : nullptr;
  // slots past width act as padding (the row's own X row, value 0), so the
  // panel loop has no branches; their gathers hit the cache
  I cofs[kChunk];
  S vi[kChunk];
  int self = -1; // a slot that gathers the row itself (the diagonal, or padding)
#pragma unroll
  for (int q = 0; q < kChunk; ++q) {
    int c = row;
    vi[q] = S(0);
    if (q < width) {
      const size_t base = size_t(q) * size_t(n);
      c = ld<true>(elem<O32>(col + base, I(row)));
      vi[q] = ld<true>(elem<O32>(val + base, I(row)));
    }
    if (c == row && self < 0) self = q;
    cofs[q] = I(c) * I(B) + I(off);
  }
  // the product's factor for this row (alpha, times the row scale when
  // given), the same for every panel
  const S* ds = static_cast<const S*>(e.dscale);
  const double arow = (AXPBY && ds) ? e.alpha * double(ld<true>(elem<O32>(ds, I(row)))) : e.alpha;
  for (unsigned p = 0; p < e.npan; ++p) {
    const size_t po = size_t(p) * e.pstride;
    const S* Xp = X + po;
    // the step's own streams (Z, F) are loaded first, so that they are in
    // flight with the gathers: one memory round trip per panel, not two
    S zr[BC] = {}, fr[BC] = {};
    if constexpr (AXPBY) {
      if (Z) RowIO<S, BC>::template load<true>(elem<O32>(Z + po, ro), zr);
      if constexpr (FORCE) RowIO<S, BC>::template load<true>(elem<O32>(F + po, ro), fr);
    }
    S x[kChunk][BC];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      I o = cofs[q];
      // O32: the 32-bit offset stays a 32-bit VGPR operand of the load (an
      // SGPR base plus a VGPR offset) instead of being widened to 64 bits
      // once outside the panel loop, which would hold two VGPRs per slot
      if constexpr (O32) asm volatile("" : "+v"(o));
      RowIO<S, BC>::load(elem<O32>(Xp, o), x[q]);
    }
    S acc[BC];
#pragma unroll
    for (int c = 0; c < BC; ++c) acc[c] = S(0);
#pragma unroll
    for (int q = 0; q < kChunk; ++q)
#pragma unroll
      for (int c = 0; c < BC; ++c) acc[c] = fma(vi[q], x[q][c], acc[c]);
    if constexpr (AXPBY) {
      // beta X needs the lane's own X row: one of the gathered rows, unless
      // the row has neither a diagonal entry nor padding
      S xr[BC];
      if (self >= 0) {
#pragma unroll
        for (int c = 0; c < BC; ++c) xr[c] = x[0][c];
#pragma unroll
        for (int q = 1; q < kChunk; ++q)
          if (q == self)
#pragma unroll
            for (int c = 0; c < BC; ++c) xr[c] = x[q][c];
      }
      else
        RowIO<S, BC>::load(elem<O32>(Xp, ro), xr);
      S out[BC];
#pragma unroll
      for (int c = 0; c < BC; ++c) {
        double v = arow * double(acc[c]) + e.beta * double(xr[c]);
        if (Z) v += e.gamma * double(zr[c]);
        if constexpr (FORCE) v += gsh[p * B + off + c] * double(fr[c]);
        out[c] = S(v);
      }
      RowIO<S, BC>::template store<true>(elem<O32>(Y + po, ro), out);
    }
    else
      epilogue<S, BC, AXPBY, O32>(acc, I(row), ro, Xp, Y + po, Z ? Z + po : nullptr, nullptr, nullptr, e);
  }
}

// A value of another lane of the wave: byte address srcaddr = 4 * source lane.
template <typename S>
__device__ inline S lane_value(int srcaddr, S v)
{
  if constexpr (sizeof(S) == 4)
    return __builtin_bit_cast(S, __builtin_amdgcn_ds_bpermute(srcaddr, __builtin_bit_cast(int, v)));
  else {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(srcaddr, int(u & 0xffffffff));
    const int hi = __builtin_amdgcn_ds_bpermute(srcaddr, int(u >> 32));
    return __builtin_bit_cast(S, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
  }
}

// The several-panel ELL step with gathers served inside the wave: the rows
// of a wave are consecutive (64 / CS of them, lanes in row order), so a
// column that falls on one of them (the +-1 neighbours of a stencil row, the
// diagonal, padding) is the X row another lane of the wave loads as its own.
// Each lane loads its own X row and only the columns outside the wave; the
// others come from the owning lane by a cross-lane permute (ds_bpermute,
// no LDS allocation), which takes those gathers off the vector memory path.
template <typename S, int B, int CS, bool AXPBY, bool O32, bool FORCE>
__global__ __launch_bounds__((panel_threads<S, B / CS>())) void k_spmm_ell_panels_sh(int width, const int32_t* __restrict__ col,
                                                                                   const S* __restrict__ val, const S* __restrict__ X, S* Y, Epi e,
                                                                                   int n)
{
  using I = Idx<O32>;
  constexpr int BC = B / CS;
  __shared__ double gsh[FORCE ? kForceCoefs : 1];
  if constexpr (FORCE) {
    for (unsigned i = threadIdx.x; i < e.npan * B; i += blockDim.x) gsh[i] = e.g[i];
    __syncthreads();
  }
  const int row = row_of(xcd_block(blockIdx.x, gridDim.x), threadIdx.x / CS, blockDim.x / CS, n, e);
  const int part = int(threadIdx.x % CS);
  const int off = part * BC;
  const bool live = row >= 0;
  // live lanes form a prefix of the wave (rows ascend with the lane until a
  // segment or the matrix ends)
  const unsigned long long liveb = __ballot(live);
  if (liveb == 0) return;
  const int nrows = int(__popcll(liveb)) / CS;
  const int r0w = __shfl(row, 0);
  const I ro = I(live ? row : r0w) * I(B) + I(off);
  const S* Z = static_cast<const S*>(e.Z);
  const S* F = FORCE ? static_cast<const S*>(e.F) : nullptr;
  int cq[kChunk];
  S vi[kChunk];
  unsigned inw = 0;  // bit q: column of slot q is a row of this wave
  unsigned anyin = 0; // bit q: some lane of the wave takes slot q from a lane
#pragma unroll
  for (int q = 0; q < kChunk; ++q) {
    int c = live ? row : r0w;
    vi[q] = S(0);
    if (live && q < width) {
      const size_t base = size_t(q) * size_t(n);
      c = ld<true>(elem<O32>(col + base, I(row)));
      vi[q] = ld<true>(elem<O32>(val + base, I(row)));
    }
    cq[q] = c;
    const bool in = live && unsigned(c - r0w) < unsigned(nrows);
    if (in) inw |= 1u << q;
    if (__ballot(in) != 0) anyin |= 1u << q;
  }
  const S* ds = static_cast<const S*>(e.dscale);
  const double arow = (AXPBY && ds && live) ? e.alpha * double(ld<true>(elem<O32>(ds, I(row)))) : e.alpha;
  for (unsigned p = 0; p < e.npan; ++p) {
    const size_t po = size_t(p) * e.pstride;
    const S* Xp = X + po;
    S zr[BC] = {}, fr[BC] = {}, xo[BC];
    // the own row first and the step's own streams (Z, F) last: the
    // permutes below wait for the own row only, with the gathers and those
    // streams still in flight (0.5-0.7% faster than Z first)
    RowIO<S, BC>::load(elem<O32>(Xp, ro), xo);
    S x[kChunk][BC];
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      if (live && !(inw >> q & 1u)) {
        I o = I(cq[q]) * I(B) + I(off);
        if constexpr (O32) asm volatile("" : "+v"(o));
        RowIO<S, BC>::load(elem<O32>(Xp, o), x[q]);
      }
    }
    if constexpr (AXPBY) {
      if (Z && live) RowIO<S, BC>::template load<true>(elem<O32>(Z + po, ro), zr);
      if constexpr (FORCE)
        if (live) RowIO<S, BC>::template load<true>(elem<O32>(F + po, ro), fr);
    }
#pragma unroll
    for (int q = 0; q < kChunk; ++q) {
      if (anyin >> q & 1u) { // wave-uniform
        const int src = ((cq[q] - r0w) * CS + part) * 4;
        const bool mine = inw >> q & 1u;
#pragma unroll
        for (int c = 0; c < BC; ++c) {
          const S t = lane_value(src, xo[c]);
          if (mine) x[q][c] = t;
        }
      }
    }
    if (!live) continue;
    S acc[BC];
#pragma unroll
    for (int c = 0; c < BC; ++c) acc[c] = S(0);
#pragma unroll
    for (int q = 0; q < kChunk; ++q)
#pragma unroll
      for (int c = 0; c < BC; ++c) acc[c] = fma(vi[q], x[q][c], acc[c]);
    S out[BC];
    if constexpr (AXPBY) {
#pragma unroll
      for (int c = 0; c < BC; ++c) {
        double v = arow * double(acc[c]) + e.beta * double(xo[c]);
        if (Z) v += e.gamma * double(zr[c]);
        if constexpr (FORCE) v += gsh[p * B + off + c] * double(fr[c]);
        out[c] = S(v);
      }
    }
    else {
      const S d = ds ? ld<true>(elem<O32>(ds, I(row))) : S(1);
#pragma unroll
      for (int c = 0; c < BC; ++c) out[c] = ds ? acc[c] * d : acc[c];
    }
    RowIO<S, BC>::template store<true>(elem<O32>(Y + po, ro), out);
  }
}

template <typename S, int B, int L, bool AXPBY, bool O32>
__global__ __launch_bounds__((spmm_threads<S, B>())) void k_spmm_lanes(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                                       const S* __restrict__ val, const S* __restrict__ X, S* Y, Epi e, int n)
{
  using I = Idx<O32>;
  const int row = row_of(xcd_block(blockIdx.x, gridDim.x), threadIdx.x / L, blockDim.x / L, n, e);
  const int sub = int(threadIdx.x) % L;
  if (row < 0) return; // whole lane groups leave together
  S acc[B];
#pragma unroll
  for (int c = 0; c < B; ++c) acc[c] = S(0);
  const int beg = rowptr[row], end = rowptr[row + 1];
  for (int p = beg + sub; p < end; p += L) {
    const int c = col[p];
    const S v = val[p];
    S x[B];
    RowIO<S, B>::load(elem<O32>(X, I(c) * I(B)), x);
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] = fma(v, x[q], acc[q]);
  }
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < B; ++q) acc[q] += __shfl_xor(acc[q], off, L);
  if (sub != 0) return;
  epilogue<S, B, AXPBY, O32>(acc, I(row), I(row) * I(B), X, Y, static_cast<const S*>(e.Z), static_cast<const S*>(e.F), e.g, e);
}

// Row-range workgroups for rows_per_wg rows each (in the blocked order, a
// whole number of workgroups per plane block), rounded up to a multiple of
// the XCD count so that the renumbering applies.
inline size_t row_ranges(size_t n, unsigned rows_per_wg, const Epi& e)
{
  size_t g;
  if (e.plane == 0)
    g = (n + rows_per_wg - 1) / rows_per_wg;
  else {
    const size_t npb = (size_t(e.plane) + e.pblock - 1) / e.pblock;
    g = npb * e.nz * ((size_t(e.pblock) + rows_per_wg - 1) / rows_per_wg);
  }
  return (g + kXcds - 1) / kXcds * kXcds;
}

// variant: 1 = one row per lane; -2 / -4 = 2 / 4 lanes per row splitting the
// columns; 2, 4, 8, 16 = lane groups splitting the nonzeros (CSR only).
// rowptr == NULL selects ELL storage (width slots).  O32: byte offsets into
// one panel and into each ELL array fit in 32 bits.
template <typename S, int B, bool AXPBY, bool O32>
int launch_rows(trl_hip_ctx* ctx, size_t n, int variant, const int32_t* rowptr, int width, const int32_t* col, const S* val, const S* X, S* Y, const Epi& e)
{
  const bool ell = rowptr == nullptr;
  const int cs = variant < 0 ? -variant : 1;
  const size_t lanes = cs > 1 ? size_t(cs) : size_t(variant > 1 ? variant : 1);
  if (variant > 1) {
    if (ell) return TRL_HIP_EINVAL;
    constexpr int T = spmm_threads<S, B>();
    const dim3 g{unsigned(row_ranges(n, T / unsigned(lanes), e))}, blk{unsigned(T)};
    switch (variant) {
    case 2:
      hipLaunchKernelGGL((k_spmm_lanes<S, B, 2, AXPBY, O32>), g, blk, 0, ctx->stream, rowptr, col, val, X, Y, e, int(n));
      break;
    case 4:
      hipLaunchKernelGGL((k_spmm_lanes<S, B, 4, AXPBY, O32>), g, blk, 0, ctx->stream, rowptr, col, val, X, Y, e, int(n));
      break;
    case 8:
      hipLaunchKernelGGL((k_spmm_lanes<S, B, 8, AXPBY, O32>), g, blk, 0, ctx->stream, rowptr, col, val, X, Y, e, int(n));
      break;
    case 16:
      hipLaunchKernelGGL((k_spmm_lanes<S, B, 16, AXPBY, O32>), g, blk, 0, ctx->stream, rowptr, col, val, X, Y, e, int(n));
      break;
    default:
      return TRL_HIP_EINVAL;
    }
    TRL_HIP_LAUNCH_CHECK();
    return 0;
  }
#define TRL_ROWS(CS_, ELL_)                                                                                                                        \
  do {                                                                                                                                             \
    constexpr int T = spmm_threads<S, B / CS_>();                                                                                                  \
    const dim3 g{unsigned(row_ranges(n, T / unsigned(lanes), e))}, blk{unsigned(T)};                                                     \
    hipLaunchKernelGGL((k_spmm_rows<S, B, CS_, ELL_, AXPBY, O32>), g, blk, 0, ctx->stream, rowptr, width, col, val, X, Y, e, int(n));            \
  } while (0)
  if (cs == 1) {
    if (ell) TRL_ROWS(1, true);
    else TRL_ROWS(1, false);
  }
  else if (cs == 2) {
    // each lane moves at least 8 bytes of a row
    if constexpr (B % 2 == 0 && (B / 2) * sizeof(S) >= 8) {
      if (ell) TRL_ROWS(2, true);
      else TRL_ROWS(2, false);
    }
    else
      return TRL_HIP_EINVAL;
  }
  else if (cs == 4) {
    if constexpr (B % 4 == 0 && (B / 4) * sizeof(S) >= 8) {
      if (ell) TRL_ROWS(4, true);
      else TRL_ROWS(4, false);
    }
    else
      return TRL_HIP_EINVAL;
  }
  else
    return TRL_HIP_EINVAL;
#undef TRL_ROWS
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

// Several panels, ELL of at most kChunk slots, row variants (1, -2, -4).
template <typename S, int B, bool AXPBY, bool O32>
int launch_ell_panels(trl_hip_ctx* ctx, size_t n, int variant, int width, const int32_t* col, const S* val, const S* X, S* Y, const Epi& e)
{
  const unsigned cs = unsigned(variant < 0 ? -variant : 1);
#define TRL_PANELS(CS_)                                                                                                                            \
  do {                                                                                                                                             \
    constexpr unsigned T = panel_threads<S, B / CS_>();                                                                                            \
    const dim3 g{unsigned(row_ranges(n, T / CS_, e))}, blk{T};                                                                                     \
    if (e.share) {                                                                                                                                 \
      if (AXPBY && e.F)                                                                                                                            \
        hipLaunchKernelGGL((k_spmm_ell_panels_sh<S, B, CS_, AXPBY, O32, AXPBY>), g, blk, 0, ctx->stream, width, col, val, X, Y, e, int(n));        \
      else                                                                                                                                         \
        hipLaunchKernelGGL((k_spmm_ell_panels_sh<S, B, CS_, AXPBY, O32, false>), g, blk, 0, ctx->stream, width, col, val, X, Y, e, int(n));        \
    }                                                                                                                                              \
    else if (AXPBY && e.F)                                                                                                                         \
      hipLaunchKernelGGL((k_spmm_ell_panels<S, B, CS_, AXPBY, O32, AXPBY>), g, blk, 0, ctx->stream, width, col, val, X, Y, e, int(n));             \
    else                                                                                                                                           \
      hipLaunchKernelGGL((k_spmm_ell_panels<S, B, CS_, AXPBY, O32, false>), g, blk, 0, ctx->stream, width, col, val, X, Y, e, int(n));             \
  } while (0)
  if (cs == 1)
    TRL_PANELS(1);
  else if (cs == 2) {
    if constexpr (B % 2 == 0 && (B / 2) * sizeof(S) >= 8)
      TRL_PANELS(2);
    else
      return TRL_HIP_EINVAL;
  }
  else if (cs == 4) {
    if constexpr (B % 4 == 0 && (B / 4) * sizeof(S) >= 8)
      TRL_PANELS(4);
    else
      return TRL_HIP_EINVAL;
  }
  else
    return TRL_HIP_EINVAL;
#undef TRL_PANELS
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

// One panel: the row or lane-group kernel.  Several panels: the ELL panel
// kernel when every row fits one chunk, else one launch per panel.
template <typename S, int B, bool AXPBY, bool O32>
int launch_panels(trl_hip_ctx* ctx, size_t n, int variant, const int32_t* rowptr, int width, const int32_t* col, const S* val, const S* X, S* Y,
                  const Epi& e)
{
  if (e.npan == 1) return launch_rows<S, B, AXPBY, O32>(ctx, n, variant, rowptr, width, col, val, X, Y, e);
  if (!rowptr && width <= kChunk && variant <= 1) {
    // a forced step stages npan * B coefficients in LDS: larger panel counts
    // go in chunks of panels
    const unsigned cap = unsigned(kForceCoefs / B);
    if (!e.F || e.npan <= cap) return launch_ell_panels<S, B, AXPBY, O32>(ctx, n, variant, width, col, val, X, Y, e);
    for (unsigned p0 = 0; p0 < e.npan; p0 += cap) {
      const size_t po = size_t(p0) * e.pstride;
      Epi ep = e;
      ep.npan = std::min(cap, e.npan - p0);
      if (ep.Z) ep.Z = static_cast<const S*>(ep.Z) + po;
      ep.F = static_cast<const S*>(ep.F) + po;
      ep.g += size_t(p0) * B;
      const int rc = launch_ell_panels<S, B, AXPBY, O32>(ctx, n, variant, width, col, val, X + po, Y + po, ep);
      if (rc) return rc;
    }
    return 0;
  }
  for (unsigned p = 0; p < e.npan; ++p) {
    const size_t po = size_t(p) * e.pstride;
    Epi ep = e;
    ep.npan = 1;
    if (ep.Z) ep.Z = static_cast<const S*>(ep.Z) + po;
    if (ep.F) {
      ep.F = static_cast<const S*>(ep.F) + po;
      ep.g += size_t(p) * B;
    }
    const int rc = launch_rows<S, B, AXPBY, O32>(ctx, n, variant, rowptr, width, col, val, X + po, Y + po, ep);
    if (rc) return rc;
  }
  return 0;
}

template <typename S, int B, bool AXPBY>
int launch_rows_any(trl_hip_ctx* ctx, size_t n, int variant, const int32_t* rowptr, int width, const int32_t* col, const S* val, const S* X, S* Y,
                    const Epi& e)
{
  const size_t lim = size_t(UINT32_MAX);
  const bool o32 = !ctx->force_offsets64 && n * size_t(B) * sizeof(S) <= lim && (rowptr || size_t(width) * n * std::max<size_t>(sizeof(S), 4) <= lim);
  if (o32) return launch_panels<S, B, AXPBY, true>(ctx, n, variant, rowptr, width, col, val, X, Y, e);
  return launch_panels<S, B, AXPBY, false>(ctx, n, variant, rowptr, width, col, val, X, Y, e);
}

// untuned default: one row per lane; FP64 rows of 8+ columns split in two
template <typename S, int B>
constexpr int default_variant()
{
  return (B * sizeof(S) >= 64) ? -2 : 1;
}

__global__ __launch_bounds__(kThreads) void k_csr_max_row(const int32_t* __restrict__ rowptr, int n, int* __restrict__ out)
{
  __shared__ int red[kThreads];
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, rowptr[i + 1] - rowptr[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < unsigned(s)) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

// max |col - row| over the matrix (CSR with sorted columns: the first and
// last entry of each row)
__global__ __launch_bounds__(kThreads) void k_csr_bandwidth(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int n, int* __restrict__ out)
{
  __shared__ int red[kThreads];
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int beg = rowptr[i], end = rowptr[i + 1];
    if (end > beg) m = max(m, max(abs(col[beg] - i), abs(col[end - 1] - i)));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < unsigned(s)) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

template <typename S>
__global__ __launch_bounds__(kThreads) void k_csr_to_ell(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, const S* __restrict__ val,
                                                         int width, int n, int32_t* __restrict__ ecol, S* __restrict__ eval)
{
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int beg = rowptr[row], len = rowptr[row + 1] - beg;
  for (int s = 0; s < width; ++s) {
    const size_t p = size_t(s) * n + row;
    const bool ok = s < len;
    ecol[p] = ok ? col[beg + s] : row;
    eval[p] = ok ? val[beg + s] : S(0);
  }
}

// max |col - row| over the entries with |col - row| < limit (the next
// bandwidth below the far off-diagonals: the line length of a grid stencil)
__global__ __launch_bounds__(kThreads) void k_csr_bandwidth_below(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col, int n, int limit,
                                                                  int* __restrict__ out)
{
  __shared__ int red[kThreads];
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int d = abs(col[p] - i);
      if (d < limit) m = max(m, d);
    }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < unsigned(s)) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

} // namespace

namespace trl_hip_impl {

int launch_csr_bandwidth_below(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, int limit, int* dev_out)
{
  hipLaunchKernelGGL(k_csr_bandwidth_below, dim3(unsigned(std::min<size_t>((n + kThreads - 1) / kThreads, 1024))), dim3(kThreads), 0, ctx->stream,
                     rowptr, col, int(n), limit, dev_out);
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_csr_max_row(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, int* dev_out)
{
  hipLaunchKernelGGL(k_csr_max_row, dim3(unsigned(std::min<size_t>((n + kThreads - 1) / kThreads, 1024))), dim3(kThreads), 0, ctx->stream, rowptr, int(n),
                     dev_out);
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

int launch_csr_to_ell(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, const void* val, int width, int32_t* ecol,
                      void* eval)
{
  const dim3 g{unsigned((n + kThreads - 1) / kThreads)}, blk{unsigned(kThreads)};
  TRL_DISPATCH_T(t, {
    hipLaunchKernelGGL((k_csr_to_ell<S>), g, blk, 0, ctx->stream, rowptr, col, static_cast<const S*>(val), width, int(n), ecol, static_cast<S*>(eval));
  });
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

bool valid_spmm_variant(int v) { return v == 0 || v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == -2 || v == -4; }

int launch_spmm(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat& A, const void* X, void* Y, const void* Z, bool axpby, double alpha,
                double beta, double gamma, int npanels, const void* F, const double* g)
{
  if ((F != nullptr) != (g != nullptr) || (F && !axpby)) return TRL_HIP_EINVAL;
  Epi e{A.dscale, Z, alpha, beta, gamma, unsigned(npanels), n * size_t(b), unsigned(A.plane), unsigned(A.plane_block), 0, F, g};
  if (A.plane > 0) e.nz = unsigned((n + size_t(A.plane) - 1) / size_t(A.plane));
  if (npanels > 1 && A.format == 1 && A.width <= kChunk) {
    e.share = unsigned(A.share);
    if (A.plane > 0 && A.tile_lines > 0 && A.line > 0) {
      // tiles of tile_lines lines: only where they cover every plane block exactly
      const int R = spmm_rows_per_block(t, b, A.variant, true);
      const int ty = A.tile_lines, ln = A.line;
      if (R > 0 && R % ty == 0 && ln % (R / ty) == 0 && A.plane_block % ln == 0 && (A.plane_block / ln) % ty == 0 && A.plane % A.plane_block == 0) {
        e.line = unsigned(ln);
        e.tly = unsigned(ty);
      }
    }
  }
  const int variant = A.variant;
  // workgroup count bound: at least 256 lanes per workgroup
  const unsigned lanes = unsigned(variant < 0 ? -variant : (variant > 1 ? variant : 1));
  if (npanels < 1 || row_ranges(n, std::max(1u, 256u / lanes), e) > size_t(UINT32_MAX)) return TRL_HIP_EINVAL;
  const int32_t* rowptr = A.format == 0 ? A.rowptr : nullptr;
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const int v = variant == 0 ? default_variant<S, B>() : variant;
      const S* vv = static_cast<const S*>(A.val);
      const S* x = static_cast<const S*>(X);
      S* y = static_cast<S*>(Y);
      if (axpby) return launch_rows_any<S, B, true>(ctx, n, v, rowptr, A.width, A.col, vv, x, y, e);
      return launch_rows_any<S, B, false>(ctx, n, v, rowptr, A.width, A.col, vv, x, y, e);
    });
  });
  return 0;
}

int spmm_rows_per_block(trl_hip_dtype t, int b, int variant, bool panels)
{
  TRL_DISPATCH_T(t, {
    TRL_DISPATCH_B(b, {
      const int v = variant == 0 ? default_variant<S, B>() : variant;
      if (v > 1) return panels ? 0 : spmm_threads<S, B>() / v;
      if (v == -2) {
        if constexpr (B % 2 == 0) return (panels ? panel_threads<S, B / 2>() : spmm_threads<S, B / 2>()) / 2;
        return 0;
      }
      if (v == -4) {
        if constexpr (B % 4 == 0) return (panels ? panel_threads<S, B / 4>() : spmm_threads<S, B / 4>()) / 4;
        return 0;
      }
      return panels ? panel_threads<S, B>() : spmm_threads<S, B>();
    });
  });
  return 0;
}

int launch_csr_bandwidth(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, int* dev_out)
{
  hipLaunchKernelGGL(k_csr_bandwidth, dim3(unsigned(std::min<size_t>((n + kThreads - 1) / kThreads, 1024))), dim3(kThreads), 0, ctx->stream, rowptr, col,
                     int(n), dev_out);
  TRL_HIP_LAUNCH_CHECK();
  return 0;
}

} // namespace trl_hip_impl
