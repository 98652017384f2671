// Internal declarations shared by the libtrl_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <unordered_map>
#include <vector>

#include "trl_hip.h"

// Context: one device, one in-order stream, and the scratch the kernels need
// for their two-level reductions.
struct trl_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 256;
  size_t max_lds = 65536; // dynamic LDS bytes a workgroup may request
  // TRL_HIP_OPT_OFFSETS64: the operator kernels take their 64-bit-offset
  // instantiations even where 32-bit lane offsets would do, and the Ritz
  // rotation its 64-bit-address form instead of its buffer-resource forms
  int force_offsets64 = 0;
  // device address of the allocation-free entry point's scratch array
  // (resolved at creation; see noalloc_scratch)
  double* noalloc_scratch = nullptr;
  // per-workgroup partial sums of the block products (FP64)
  double* partials = nullptr;
  size_t partials_cap = 0; // in doubles
  // bumped whenever a scratch buffer is reallocated: a captured graph that
  // names the old one is stale
  unsigned long long scratch_epoch = 0;
  // stream capture in progress (trl_hip_graph_begin .. _end); a call that
  // needed the host was refused during it
  int capturing = 0, capture_refused = 0;
  // TRL_HIP_OPT_GRAPHS: drivers replay captured restart cycles
  int graphs = 0;
  // device allocations of trl_hip_malloc (TRL_HIP_OPT_ALLOC_CACHE): the sizes
  // of live blocks, and freed blocks kept for reuse by size (their reuse is
  // ordered on the context stream, so a free needs no synchronisation)
  int alloc_cache = 1;
  std::unordered_map<void*, size_t> live_sizes;
  std::multimap<size_t, void*> cached;
  size_t cached_bytes = 0;
  // flags[0]: sticky Cholesky breakdown pivot (1-based, 0 = none); [8]:
  // CholQR arrival counter; [16]: csr_max_row result; [24]: second-pass gate
  int* flags = nullptr;
  // small FP64 scratch: R^{-1}, R of the first CholQR pass, coefficient staging
  double* small = nullptr;
  // coefficient scratch for multi-block products (grown on demand)
  double* coef = nullptr;
  size_t coef_cap = 0;
  // profiling mode: HIP events around every queued kernel family call
  struct ProfRec {
    hipEvent_t start, stop;
    int family;
    double bytes;
    int dtype;
  };
  int prof_on = 0;
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  long long prof_count[TRL_HIP_PROF_FAMILIES] = {};
  double prof_seconds[TRL_HIP_PROF_FAMILIES] = {};
  double prof_bytes[TRL_HIP_PROF_FAMILIES] = {};
  // the same split by panel dtype (index trl_hip_dtype)
  long long prof_t_count[TRL_HIP_PROF_FAMILIES][2] = {};
  double prof_t_seconds[TRL_HIP_PROF_FAMILIES][2] = {};
  double prof_t_bytes[TRL_HIP_PROF_FAMILIES][2] = {};
  // profiling mode: trl_hip_reorth calls and how many took the second pass
  long long reorth_calls = 0, reorth_second = 0;
  // region timer events (created on first use)
  hipEvent_t region[2] = {nullptr, nullptr};
  std::unordered_map<const void*, long long> nnz_cache;
};

namespace trl_hip_impl {

constexpr int kSmallScratch = 4 * 256; // doubles in ctx->small

// Grows ctx->partials to hold at least `doubles` entries.
int ensure_partials(trl_hip_ctx* ctx, size_t doubles);

inline bool valid_b(int b) { return b == 1 || b == 2 || b == 4 || b == 6 || b == 8 || b == 16; }

// launch wrappers (one translation unit each family)
int launch_fill_zero(trl_hip_ctx* ctx, void* dst, size_t bytes);
int launch_convert(trl_hip_ctx* ctx, trl_hip_dtype st, const void* src, trl_hip_dtype dt, void* dst, size_t count);
int launch_panel_sub(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const void* Y);
int launch_panel_mult(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const double* W, void* Y, int mode);
// gate (device int, may be NULL): the kernels return at once when *gate == 0
// partials/cap: a fixed buffer of cap doubles for the per-workgroup partial
// sums instead of the context's growable one (the workgroup count is capped
// to fit; TRL_HIP_EINVAL if one partial row per workgroup does not fit)
int launch_gram(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const void* W, const void* md, double* C,
                const int* gate = nullptr, double* partials = nullptr, size_t cap = 0);
// Scratch of the allocation-free Gram entry point: kNoAllocScratchDoubles
// doubles per device, resolved once per device (the device must be current
// on the first call for it).
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 20;
int noalloc_scratch(int device, void** out);
int launch_update(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const double* C, void* W, const int* gate = nullptr);
int launch_dgks(trl_hip_ctx* ctx, int b, int k, const double* C, double eta2, int* gate);
int launch_cholqr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, const void* md, double* R, int passes);
// CholQR in the B-inner product with BW = B W given; BW is scaled with W
int launch_cholqr_bprod(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, void* BW, double* R, int passes);
// the shift-invert inner solver (FP64): workspace bytes, and the solve
size_t pcg_workspace_bytes(size_t n, int b);
int launch_pcg(trl_hip_ctx* ctx, size_t n, int b, const trl_hip_spmat& K, const double* dinv, const double* rhs, double* X, double tol, int maxit,
               void* work, int* iters, double* relres);
int launch_rotate(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int m, const void* V, const double* Y, int ybc, int kc, void* Out);
bool valid_spmm_variant(int v);
// F, g (both or neither, axpby only): the forcing term + F_p diag(g_p), F
// laid out as X, g device doubles, b per panel
int launch_spmm(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat& A, const void* X, void* Y, const void* Z, bool axpby, double alpha,
                double beta, double gamma, int npanels = 1, const void* F = nullptr, const double* g = nullptr);
// rows per workgroup of the one-panel (panels = false) or several-panel ELL
// product kernel of this variant
int spmm_rows_per_block(trl_hip_dtype t, int b, int variant, bool panels = false);
int launch_csr_bandwidth(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, int* dev_out);
// max |col - row| over the entries with |col - row| < limit
int launch_csr_bandwidth_below(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, int limit, int* dev_out);
int launch_csr_max_row(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, int* dev_out);
int launch_csr_to_ell(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, const void* val, int width, int32_t* ecol,
                      void* eval);
int launch_small(trl_hip_ctx* ctx, int op, int b, const double* A, const double* B, double* C, const double* hostvals);
int launch_residual_split(trl_hip_ctx* ctx, size_t n, int b, int npanels, const double* AX, const double* X, const double* theta, double s, float* R,
                          float* E);
int launch_combine(trl_hip_ctx* ctx, size_t n, int b, int npanels, double* X, const double* h, const float* E);
int launch_residual_sq(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const void* X, const void* Y, const void* md, const double* lambda, double* out);

inline int hip_status(hipError_t e) { return e == hipSuccess ? 0 : static_cast<int>(e); }

// A call that synchronises with the host, allocates or reads device memory
// back cannot be part of a captured graph: while the context stream is being
// captured it is refused before it touches the stream (so the capture stays
// valid, and trl_hip_graph_end discards it).
inline int refuse_in_capture(trl_hip_ctx* ctx)
{
  if (!ctx->capturing) return 0;
  ctx->capture_refused = 1;
  return TRL_HIP_EINVAL;
}
#define TRL_HIP_NO_CAPTURE(ctx)                                                                                                                       \
  do {                                                                                                                                                 \
    if (const int rc_capture_ = refuse_in_capture(ctx)) return rc_capture_;                                                                           \
  } while (0)

} // namespace trl_hip_impl

#define TRL_HIP_TRY(expr)                                                                                                                            \
  do {                                                                                                                                               \
    hipError_t e_ = (expr);                                                                                                                          \
    if (e_ != hipSuccess) return static_cast<int>(e_);                                                                                               \
  } while (0)

#define TRL_HIP_LAUNCH_CHECK() TRL_HIP_TRY(hipGetLastError())

// Runtime block size -> compile-time constant B.
#define TRL_DISPATCH_B(bval, ...)                                                                                                                    \
  switch (bval) {                                                                                                                                    \
  case 1: {                                                                                                                                          \
    constexpr int B = 1;                                                                                                                             \
    __VA_ARGS__;                                                                                                                                     \
  } break;                                                                                                                                           \
  case 2: {                                                                                                                                          \
    constexpr int B = 2;                                                                                                                             \
    __VA_ARGS__;                                                                                                                                     \
  } break;                                                                                                                                           \
  case 4: {                                                                                                                                          \
    constexpr int B = 4;                                                                                                                             \
    __VA_ARGS__;                                                                                                                                     \
  } break;                                                                                                                                           \
  case 6: {                                                                                                                                          \
    constexpr int B = 6;                                                                                                                             \
    __VA_ARGS__;                                                                                                                                     \
  } break;                                                                                                                                           \
  case 8: {                                                                                                                                          \
    constexpr int B = 8;                                                                                                                             \
    __VA_ARGS__;                                                                                                                                     \
  } break;                                                                                                                                           \
  case 16: {                                                                                                                                         \
    constexpr int B = 16;                                                                                                                            \
    __VA_ARGS__;                                                                                                                                     \
  } break;                                                                                                                                           \
  default:                                                                                                                                           \
    return TRL_HIP_EINVAL;                                                                                                                           \
  }

#define TRL_DISPATCH_T(tval, ...)                                                                                                                    \
  if ((tval) == TRL_HIP_F32) {                                                                                                                       \
    using S = float;                                                                                                                                 \
    __VA_ARGS__;                                                                                                                                     \
  }                                                                                                                                                  \
  else if ((tval) == TRL_HIP_F64) {                                                                                                                  \
    using S = double;                                                                                                                                \
    __VA_ARGS__;                                                                                                                                     \
  }                                                                                                                                                  \
  else                                                                                                                                               \
    return TRL_HIP_EINVAL;

// ---------------------------------------------------------------------------
// device helpers
namespace trl_hip_dev {

// Plain or non-temporal (streaming: read or written once, kept out of the
// way of reused data in the caches) element access.
template <bool NT, typename T>
__device__ inline T ld(const T* p)
{
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool NT, typename T>
__device__ inline void st(T* p, const T& v)
{
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// Loads/stores of one panel row (B contiguous scalars) with the widest
// aligned vector the row allows (rows start at r*B*sizeof(S)); NT selects
// non-temporal accesses.
template <typename S, int B>
struct RowIO {
  static constexpr int bytes = B * int(sizeof(S));
  static constexpr int vec = (bytes % 16 == 0) ? 16 : (bytes % 8 == 0) ? 8 : int(sizeof(S));
  static constexpr int per = vec / int(sizeof(S)); // scalars per vector
  static constexpr int nv = B / per;
  using V16 = __attribute__((ext_vector_type(16 / sizeof(S)))) S;

  template <bool NT = false>
  __device__ static inline void load(const S* __restrict__ p, S (&v)[B])
  {
    if constexpr (vec == 16) {
      using VT = __attribute__((ext_vector_type(16 / sizeof(S)))) S;
      const VT* q = reinterpret_cast<const VT*>(p);
#pragma unroll
      for (int i = 0; i < nv; ++i) {
        VT t = ld<NT>(q + i);
#pragma unroll
        for (int e = 0; e < per; ++e) v[i * per + e] = t[e];
      }
    }
    else if constexpr (vec == 8 && per == 2) {
      using VT = __attribute__((ext_vector_type(2))) S;
      const VT* q = reinterpret_cast<const VT*>(p);
#pragma unroll
      for (int i = 0; i < nv; ++i) {
        VT t = ld<NT>(q + i);
        v[2 * i] = t[0];
        v[2 * i + 1] = t[1];
      }
    }
    else {
#pragma unroll
      for (int i = 0; i < B; ++i) v[i] = ld<NT>(p + i);
    }
  }

  template <bool NT = false>
  __device__ static inline void store(S* __restrict__ p, const S (&v)[B])
  {
    if constexpr (vec == 16) {
      using VT = __attribute__((ext_vector_type(16 / sizeof(S)))) S;
      VT* q = reinterpret_cast<VT*>(p);
#pragma unroll
      for (int i = 0; i < nv; ++i) {
        VT t;
#pragma unroll
        for (int e = 0; e < per; ++e) t[e] = v[i * per + e];
        st<NT>(q + i, t);
      }
    }
    else if constexpr (vec == 8 && per == 2) {
      using VT = __attribute__((ext_vector_type(2))) S;
      VT* q = reinterpret_cast<VT*>(p);
#pragma unroll
      for (int i = 0; i < nv; ++i) {
        VT t;
        t[0] = v[2 * i];
        t[1] = v[2 * i + 1];
        st<NT>(q + i, t);
      }
    }
    else {
#pragma unroll
      for (int i = 0; i < B; ++i) st<NT>(p + i, v[i]);
    }
  }
};

// Error-free transformation a + b = s + e (Knuth's TwoSum).
template <typename T>
__device__ inline void two_sum(T a, T b, T& s, T& e)
{
  s = a + b;
  T bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Range-checked buffer accesses (raw buffer resource, no stride): an access
// at a byte offset at or past the resource's size reads zeros and writes
// nothing, so edge lanes need no branch around their loads and stores and
// the compiler can count the loads and stores a wave has in flight exactly.
constexpr int kBufFlags = 0x00020000;     // resource word 3 on gfx9-class parts

template <int BYTES>
struct BufIO;
template <>
struct BufIO<16> {
  using T = unsigned __attribute__((ext_vector_type(4)));
  template <int AUX = 0>
  __device__ static T load(__amdgpu_buffer_rsrc_t r, unsigned off)
  {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
  }
  template <int AUX = 0>
  __device__ static void store(T v, __amdgpu_buffer_rsrc_t r, unsigned off)
  {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, AUX);
  }
};
template <>
struct BufIO<8> {
  using T = unsigned __attribute__((ext_vector_type(2)));
  template <int AUX = 0>
  __device__ static T load(__amdgpu_buffer_rsrc_t r, unsigned off)
  {
    return __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX);
  }
  template <int AUX = 0>
  __device__ static void store(T v, __amdgpu_buffer_rsrc_t r, unsigned off)
  {
    __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, AUX);
  }
};
template <>
struct BufIO<4> {
  using T = unsigned;
  template <int AUX = 0>
  __device__ static T load(__amdgpu_buffer_rsrc_t r, unsigned off)
  {
    return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX);
  }
  template <int AUX = 0>
  __device__ static void store(T v, __amdgpu_buffer_rsrc_t r, unsigned off)
  {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, AUX);
  }
};

template <typename T>
__device__ inline T buf_load(__amdgpu_buffer_rsrc_t r, unsigned off)
{
  return __builtin_bit_cast(T, BufIO<sizeof(T)>::load(r, off));
}
template <typename T>
__device__ inline void buf_store(const T& v, __amdgpu_buffer_rsrc_t r, unsigned off)
{
  BufIO<sizeof(T)>::store(__builtin_bit_cast(typename BufIO<sizeof(T)>::T, v), r, off);
}
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, size_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), short(0), int(bytes), kBufFlags);
}
// The same for a base and size that are the same in every lane but computed
// where the compiler cannot tell: read from the first lane, so the resource
// sits in scalar registers (no per-lane loop around the access).
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc_uniform(const void* p, size_t bytes)
{
  const unsigned long long a = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(a)), hi = __builtin_amdgcn_readfirstlane(unsigned(a >> 32));
  const int nb = __builtin_amdgcn_readfirstlane(int(bytes));
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>((static_cast<unsigned long long>(hi) << 32) | lo), short(0), nb, kBufFlags);
}

} // namespace trl_hip_dev
