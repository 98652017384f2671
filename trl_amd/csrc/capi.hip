// C-ABI entry points of libtrl_hip.so: argument checks, context and memory
// management, dispatch to the kernel families.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "common.hh"

using namespace trl_hip_impl;

namespace trl_hip_impl {

int ensure_partials(trl_hip_ctx* ctx, size_t doubles)
{
  if (doubles <= ctx->partials_cap) return 0;
  TRL_HIP_NO_CAPTURE(ctx);
  size_t cap = ctx->partials_cap ? ctx->partials_cap : (size_t(1) << 16);
  while (cap < doubles) cap *= 2;
  // the stream may still use the old buffer
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (ctx->partials) TRL_HIP_TRY(hipFree(ctx->partials));
  ctx->partials = nullptr;
  ctx->partials_cap = 0;
  if (hipMalloc(&ctx->partials, cap * sizeof(double)) != hipSuccess) return TRL_HIP_ENOMEM;
  ctx->partials_cap = cap;
  ++ctx->scratch_epoch;
  return 0;
}

// RAII event pair around one queued call (profiling mode only).
struct Prof {
  trl_hip_ctx* ctx;
  int family;
  double bytes;
  int dtype; // trl_hip_dtype of the call, -1 if none
  hipEvent_t start = nullptr;
  Prof(trl_hip_ctx* c, int fam, double b, int t = -1) : ctx(c), family(fam), bytes(b), dtype(t)
  {
    if (!ctx->prof_on) return;
    start = take();
    if (start) (void)hipEventRecord(start, ctx->stream);
  }
  hipEvent_t take()
  {
    if (!ctx->prof_pool.empty()) {
      hipEvent_t e = ctx->prof_pool.back();
      ctx->prof_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  ~Prof()
  {
    if (!ctx->prof_on || !start) return;
    hipEvent_t stop = take();
    if (!stop) return;
    (void)hipEventRecord(stop, ctx->stream);
    ctx->prof_pending.push_back({start, stop, family, bytes, dtype});
  }
};

int prof_fold(trl_hip_ctx* ctx)
{
  if (ctx->prof_pending.empty()) return 0;
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->prof_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess) {
      ctx->prof_seconds[r.family] += 1e-3 * ms;
      ctx->prof_bytes[r.family] += r.bytes;
      ctx->prof_count[r.family] += 1;
      if (r.dtype == TRL_HIP_F32 || r.dtype == TRL_HIP_F64) {
        ctx->prof_t_seconds[r.family][r.dtype] += 1e-3 * ms;
        ctx->prof_t_bytes[r.family][r.dtype] += r.bytes;
        ctx->prof_t_count[r.family][r.dtype] += 1;
      }
    }
    ctx->prof_pool.push_back(r.start);
    ctx->prof_pool.push_back(r.stop);
  }
  ctx->prof_pending.clear();
  return 0;
}

double panel_bytes(trl_hip_dtype t, size_t n, int b) { return double(n) * b * (t == TRL_HIP_F32 ? 4.0 : 8.0); }

long long csr_nnz(trl_hip_ctx* ctx, const int32_t* rowptr, size_t n)
{
  auto it = ctx->nnz_cache.find(rowptr);
  if (it != ctx->nnz_cache.end()) return it->second;
  if (refuse_in_capture(ctx)) return 0;
  int32_t last = 0;
  if (hipMemcpyAsync(&last, rowptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) return 0;
  (void)hipStreamSynchronize(ctx->stream);
  ctx->nnz_cache[rowptr] = last;
  return last;
}

} // namespace trl_hip_impl

extern "C" {

const char* trl_hip_strerror(int code)
{
  switch (code) {
  case TRL_HIP_OK:
    return "success";
  case TRL_HIP_EINVAL:
    return "invalid argument";
  case TRL_HIP_ECHOL:
    return "Cholesky factorization failed in orthonormalize";
  case TRL_HIP_ENOMEM:
    return "device allocation failed";
  case TRL_HIP_EEIG:
    return "Eigendecomposition failed";
  case TRL_HIP_ERANGE:
    return "block index out of range";
  case TRL_HIP_EBARRIER:
    return "device barrier of a batched step timed out";
  case TRL_HIP_ENOCONV:
    return "inner solve did not converge";
  case TRL_HIP_EINDEF:
    return "shifted operator is not positive definite";
  case TRL_HIP_ESTALE:
    return "graph names scratch memory the context has since replaced";
  case TRL_HIP_EIO:
    return "cannot open or parse the file";
  default:
    if (code > 0) return hipGetErrorString(static_cast<hipError_t>(code));
    return "unknown trl_hip error";
  }
}

int trl_hip_ctx_create(int device, trl_hip_ctx** out)
{
  if (!out) return TRL_HIP_EINVAL;
  *out = nullptr;
  int ndev = 0;
  TRL_HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return TRL_HIP_EINVAL;
  TRL_HIP_TRY(hipSetDevice(device));
  auto* ctx = new (std::nothrow) trl_hip_ctx();
  if (!ctx) return TRL_HIP_ENOMEM;
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->num_cus = prop.multiProcessorCount;
    if (prop.sharedMemPerBlock > 0) ctx->max_lds = prop.sharedMemPerBlock;
  }
  // default-priority, non-blocking in-order stream
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  // cleared on the context stream itself, so that its first kernels see zeros
  if (e == hipSuccess) e = hipMalloc(&ctx->flags, 64 * sizeof(int));
  if (e == hipSuccess) e = hipMemsetAsync(ctx->flags, 0, 64 * sizeof(int), ctx->stream);
  if (e == hipSuccess) e = hipMalloc(&ctx->small, kSmallScratch * sizeof(double));
  if (e == hipSuccess) e = hipMemsetAsync(ctx->small, 0, kSmallScratch * sizeof(double), ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    trl_hip_ctx_destroy(ctx);
    return static_cast<int>(e);
  }
  int rc = ensure_partials(ctx, size_t(1) << 18);
  // the allocation-free entry point's scratch on this device (current here)
  void* scr = nullptr;
  if (!rc) rc = noalloc_scratch(device, &scr);
  if (rc) {
    trl_hip_ctx_destroy(ctx);
    return rc;
  }
  ctx->noalloc_scratch = static_cast<double*>(scr);
  *out = ctx;
  return 0;
}

namespace {

// bytes of freed blocks a context keeps for reuse at most
constexpr size_t kAllocCacheCap = size_t(64) << 30;

// Releases the cached blocks (after the work queued on them has finished).
void flush_alloc_cache(trl_hip_ctx* ctx)
{
  if (ctx->cached.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->cached) (void)hipFree(kv.second);
  ctx->cached.clear();
  ctx->cached_bytes = 0;
}

} // namespace

int trl_hip_ctx_destroy(trl_hip_ctx* ctx)
{
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  flush_alloc_cache(ctx);
  if (ctx->partials) (void)hipFree(ctx->partials);
  if (ctx->flags) (void)hipFree(ctx->flags);
  if (ctx->small) (void)hipFree(ctx->small);
  if (ctx->coef) (void)hipFree(ctx->coef);
  for (auto& r : ctx->prof_pending) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : ctx->prof_pool) (void)hipEventDestroy(e);
  for (auto e : ctx->region)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int trl_hip_ctx_sync(trl_hip_ctx* ctx)
{
  if (!ctx) return TRL_HIP_EINVAL;
  TRL_HIP_NO_CAPTURE(ctx);
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

int trl_hip_ctx_check(trl_hip_ctx* ctx, int* pivot)
{
  if (!ctx) return TRL_HIP_EINVAL;
  TRL_HIP_NO_CAPTURE(ctx);
  int flag = 0;
  TRL_HIP_TRY(hipMemcpyAsync(&flag, ctx->flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (pivot) *pivot = flag;
  if (flag != 0) {
    int zero = 0;
    TRL_HIP_TRY(hipMemcpyAsync(ctx->flags, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return TRL_HIP_ECHOL;
  }
  return 0;
}

void* trl_hip_ctx_stream(trl_hip_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int trl_hip_ctx_num_cus(trl_hip_ctx* ctx) { return ctx ? ctx->num_cus : 0; }

int trl_hip_ctx_set_option(trl_hip_ctx* ctx, int option, int value)
{
  if (!ctx) return TRL_HIP_EINVAL;
  switch (option) {
  case TRL_HIP_OPT_OFFSETS64:
    ctx->force_offsets64 = value ? 1 : 0;
    return 0;
  case TRL_HIP_OPT_GRAPHS:
    ctx->graphs = value ? 1 : 0;
    return 0;
  case TRL_HIP_OPT_ALLOC_CACHE:
    TRL_HIP_NO_CAPTURE(ctx);
    ctx->alloc_cache = value ? 1 : 0;
    if (!value) flush_alloc_cache(ctx);
    return 0;

  default:
    return TRL_HIP_EINVAL;
  }
}

int trl_hip_ctx_get_option(trl_hip_ctx* ctx, int option, int* value)
{
  if (!ctx || !value) return TRL_HIP_EINVAL;
  switch (option) {
  case TRL_HIP_OPT_OFFSETS64:
    *value = ctx->force_offsets64;
    return 0;
  case TRL_HIP_OPT_GRAPHS:
    *value = ctx->graphs;
    return 0;
  case TRL_HIP_OPT_ALLOC_CACHE:
    *value = ctx->alloc_cache;
    return 0;
  default:
    return TRL_HIP_EINVAL;
  }
}


int trl_hip_malloc(trl_hip_ctx* ctx, size_t bytes, void** out)
{
  if (!ctx || !out) return TRL_HIP_EINVAL;
  *out = nullptr;
  TRL_HIP_NO_CAPTURE(ctx);
  if (bytes == 0) bytes = 1;
  const size_t aligned = (bytes + 255) & ~size_t(255);
  void* p = nullptr;
  // a freed block of the same size, reused in stream order (a hipMalloc of
  // hundreds of MB and its hipFree cost ~0.05 and ~0.3 ms of host time, and
  // the free a synchronisation, inside every solve that allocates scratch)
  if (ctx->alloc_cache) {
    auto it = ctx->cached.find(aligned);
    if (it != ctx->cached.end()) {
      p = it->second;
      ctx->cached.erase(it);
      ctx->cached_bytes -= aligned;
    }
  }
  if (!p && hipMalloc(&p, aligned) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    flush_alloc_cache(ctx); // the cached blocks may be what is missing
    if (hipMalloc(&p, aligned) != hipSuccess) {
      (void)hipGetLastError();
      return TRL_HIP_ENOMEM;
    }
  }
  int rc = launch_fill_zero(ctx, p, aligned);
  if (rc) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(p);
    return rc;
  }
  ctx->live_sizes[p] = aligned;
  *out = p;
  return 0;
}

int trl_hip_free(trl_hip_ctx* ctx, void* p)
{
  if (!ctx) return TRL_HIP_EINVAL;
  if (!p) return 0;
  TRL_HIP_NO_CAPTURE(ctx);
  auto it = ctx->live_sizes.find(p);
  if (it != ctx->live_sizes.end()) {
    const size_t sz = it->second;
    ctx->live_sizes.erase(it);
    if (ctx->alloc_cache && ctx->cached_bytes + sz <= kAllocCacheCap) {
      ctx->cached.emplace(sz, p);
      ctx->cached_bytes += sz;
      return 0;
    }
  }
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  TRL_HIP_TRY(hipFree(p));
  return 0;
}

int trl_hip_memcpy_h2d(trl_hip_ctx* ctx, void* dst, const void* src, size_t bytes)
{
  if (!ctx || (bytes && (!dst || !src))) return TRL_HIP_EINVAL;
  if (!bytes) return 0;
  TRL_HIP_NO_CAPTURE(ctx);
  TRL_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

int trl_hip_memcpy_d2h(trl_hip_ctx* ctx, void* dst, const void* src, size_t bytes)
{
  if (!ctx || (bytes && (!dst || !src))) return TRL_HIP_EINVAL;
  if (!bytes) return 0;
  TRL_HIP_NO_CAPTURE(ctx);
  TRL_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

int trl_hip_memcpy_d2d(trl_hip_ctx* ctx, void* dst, const void* src, size_t bytes)
{
  if (!ctx || (bytes && (!dst || !src))) return TRL_HIP_EINVAL;
  if (!bytes || dst == src) return 0;
  TRL_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}

int trl_hip_fill_zero(trl_hip_ctx* ctx, void* dst, size_t bytes)
{
  if (!ctx || (bytes && !dst)) return TRL_HIP_EINVAL;
  if (!bytes) return 0;
  return launch_fill_zero(ctx, dst, bytes);
}

int trl_hip_convert(trl_hip_ctx* ctx, trl_hip_dtype st, const void* src, trl_hip_dtype dt, void* dst, size_t count)
{
  if (!ctx || (count && (!src || !dst))) return TRL_HIP_EINVAL;
  if (!count) return 0;
  return launch_convert(ctx, st, src, dt, dst, count);
}

int trl_hip_panel_sub(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const void* Y)
{
  if (!ctx || !valid_b(b) || (n && (!X || !Y))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_PANEL, 3 * panel_bytes(t, n, b), t);
  return launch_panel_sub(ctx, t, n, b, X, Y);
}

int trl_hip_panel_mult(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* X, const double* W, void* Y, int mode)
{
  if (!ctx || !valid_b(b) || mode < 0 || mode > 4 || (n && (!X || !W || !Y))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_PANEL, ((mode == 1 || mode >= 3) ? 3 : 2) * panel_bytes(t, n, b), t);
  return launch_panel_mult(ctx, t, n, b, X, W, Y, mode);
}

int trl_hip_gram(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const void* W, const void* mdiag, double* C)
{
  if (!ctx || !valid_b(b) || k < 1 || !C || (n && (!V || !W))) return TRL_HIP_EINVAL;
  Prof pr(ctx, TRL_HIP_PROF_GRAM, (k + 1) * panel_bytes(t, n, b) + (mdiag ? panel_bytes(t, n, 1) : 0.0), t);
  return launch_gram(ctx, t, n, b, k, V, W, mdiag, C);
}

int trl_hip_gram_noalloc(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const void* W, const void* mdiag, double* C)
{
  if (!ctx || !valid_b(b) || k < 1 || !C || (n && (!V || !W)) || !ctx->noalloc_scratch) return TRL_HIP_EINVAL;
  if (size_t(k) * b * b > TRL_HIP_NOALLOC_MAX_COEFS) return TRL_HIP_EINVAL;
  // the same kernels as trl_hip_gram, their partial sums in the device's
  // fixed scratch array (no profiling events: those are created on demand)
  return launch_gram(ctx, t, n, b, k, V, W, mdiag, C, nullptr, ctx->noalloc_scratch, kNoAllocScratchDoubles);
}

int trl_hip_update(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, const double* C, void* W)
{
  if (!ctx || !valid_b(b) || k < 1 || !C || (n && (!V || !W))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_UPDATE, (k + 2) * panel_bytes(t, n, b), t);
  return launch_update(ctx, t, n, b, k, V, C, W);
}

int trl_hip_reorth(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int k, const void* V, void* W, const void* mdiag, double* C, double eta)
{
  if (!ctx || !valid_b(b) || k < 0 || !C || (n && (!V || !W)) || !(eta < 1.0)) return TRL_HIP_EINVAL;
  if (k == 0 || n == 0) return 0;
  const size_t BB = size_t(b) * b;
  double* C1 = C;
  double* C2 = C + (k + 1) * BB;
  const double P = panel_bytes(t, n, b), Pm = mdiag ? panel_bytes(t, n, 1) : 0.0;
  int rc = 0;
  if (eta <= 0.0) { // two full passes
    for (double* Cp : {C1, C2}) {
      {
        Prof pr(ctx, TRL_HIP_PROF_GRAM, (k + 1) * P + Pm, t);
        if ((rc = launch_gram(ctx, t, n, b, k, V, W, mdiag, Cp))) return rc;
      }
      Prof pr(ctx, TRL_HIP_PROF_UPDATE, (k + 2) * P, t);
      if ((rc = launch_update(ctx, t, n, b, k, V, Cp, W))) return rc;
    }
    return 0;
  }
  const size_t es = (t == TRL_HIP_F32) ? 4 : 8;
  const bool contiguous = static_cast<const char*>(V) + size_t(k) * n * b * es == static_cast<const char*>(W);
  {
    // pass 1 with W^T M W in the same sweep when W follows the k blocks
    Prof pr(ctx, TRL_HIP_PROF_GRAM, (k + 1) * P + Pm + (contiguous ? 0.0 : 2 * P + Pm), t);
    if (contiguous)
      rc = launch_gram(ctx, t, n, b, k + 1, V, W, mdiag, C1);
    else {
      rc = launch_gram(ctx, t, n, b, k, V, W, mdiag, C1);
      if (!rc) rc = launch_gram(ctx, t, n, b, 1, W, W, mdiag, C1 + k * BB);
    }
    if (rc) return rc;
  }
  int* gate = ctx->flags + 24;
  if ((rc = launch_dgks(ctx, b, k, C1, eta * eta, gate))) return rc;
  {
    Prof pr(ctx, TRL_HIP_PROF_UPDATE, (k + 2) * P, t);
    if ((rc = launch_update(ctx, t, n, b, k, V, C1, W))) return rc;
  }
  if (ctx->prof_on) {
    // profiling mode: decide on the host so that skipped work is not counted
    TRL_HIP_NO_CAPTURE(ctx);
    int g = 1;
    TRL_HIP_TRY(hipMemcpyAsync(&g, gate, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->reorth_calls += 1;
    if (!g) return 0;
    ctx->reorth_second += 1;
    {
      Prof pr(ctx, TRL_HIP_PROF_GRAM, (k + 1) * P + Pm, t);
      if ((rc = launch_gram(ctx, t, n, b, k, V, W, mdiag, C2))) return rc;
    }
    Prof pr(ctx, TRL_HIP_PROF_UPDATE, (k + 2) * P, t);
    return launch_update(ctx, t, n, b, k, V, C2, W);
  }
  if ((rc = launch_gram(ctx, t, n, b, k, V, W, mdiag, C2, gate))) return rc;
  return launch_update(ctx, t, n, b, k, V, C2, W, gate);
}

int trl_hip_cholqr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, const void* mdiag, double* R, int passes)
{
  if (!ctx || !valid_b(b) || !R || passes < 1 || passes > 2 || (n && !W)) return TRL_HIP_EINVAL;
  Prof pr(ctx, TRL_HIP_PROF_CHOLQR, passes * (3 * panel_bytes(t, n, b) + (mdiag ? panel_bytes(t, n, 1) : 0.0)), t);
  return launch_cholqr(ctx, t, n, b, W, mdiag, R, passes);
}

} // extern "C"

// A captured, instantiated graph and the context state it depends on.
struct trl_hip_graph {
  hipGraphExec_t exec = nullptr;
  const trl_hip_ctx* ctx = nullptr;
  unsigned long long epoch = 0;
};

extern "C" {

int trl_hip_graph_begin(trl_hip_ctx* ctx)
{
  // profiling mode records events around every call and reads them later
  if (!ctx || ctx->capturing || ctx->prof_on) return TRL_HIP_EINVAL;
  TRL_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = 1;
  ctx->capture_refused = 0;
  return 0;
}

int trl_hip_graph_end(trl_hip_ctx* ctx, trl_hip_graph** graph)
{
  if (!ctx || !graph || !ctx->capturing) return TRL_HIP_EINVAL;
  *graph = nullptr;
  ctx->capturing = 0;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  (void)hipGetLastError(); // a call refused during the capture leaves its error behind
  if (e != hipSuccess || !g || ctx->capture_refused) {
    // a call refused during the capture: the graph lacks its work
    if (g) (void)hipGraphDestroy(g);
    ctx->capture_refused = 0;
    return e != hipSuccess ? int(e) : TRL_HIP_EINVAL;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return int(ei);
  auto* out = new (std::nothrow) trl_hip_graph;
  if (!out) {
    (void)hipGraphExecDestroy(exec);
    return TRL_HIP_ENOMEM;
  }
  out->exec = exec;
  out->ctx = ctx;
  out->epoch = ctx->scratch_epoch;
  *graph = out;
  return 0;
}

int trl_hip_graph_launch(trl_hip_ctx* ctx, const trl_hip_graph* graph)
{
  if (!ctx || !graph || !graph->exec || graph->ctx != ctx || ctx->capturing) return TRL_HIP_EINVAL;
  if (graph->epoch != ctx->scratch_epoch) return TRL_HIP_ESTALE;
  TRL_HIP_TRY(hipGraphLaunch(graph->exec, ctx->stream));
  return 0;
}

int trl_hip_graph_destroy(trl_hip_graph* graph)
{
  if (!graph) return 0;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  delete graph;
  return 0;
}

int trl_hip_cholqr_bprod(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, void* W, void* BW, double* R, int passes)
{
  if (!ctx || !valid_b(b) || !R || passes < 1 || passes > 2 || (n && (!W || !BW)) || (n && W == BW)) return TRL_HIP_EINVAL;
  Prof pr(ctx, TRL_HIP_PROF_CHOLQR, passes * 6 * panel_bytes(t, n, b), t);
  return launch_cholqr_bprod(ctx, t, n, b, W, BW, R, passes);
}

int trl_hip_rotate(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int m, const void* V, const double* Y, int y_block_cols, int kc, void* Out)
{
  if (!ctx || !valid_b(b) || m < 1 || kc < 1 || y_block_cols < kc || !Y || (n && (!V || !Out))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_ROTATE, (m + kc) * panel_bytes(t, n, b), t);
  return launch_rotate(ctx, t, n, b, m, V, Y, y_block_cols, kc, Out);
}

namespace {

double spmm_bytes(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, int panels)
{
  const double s = (t == TRL_HIP_F32) ? 4.0 : 8.0;
  double bytes = panels * panel_bytes(t, n, b) + (A->dscale ? double(n) * s : 0.0);
  if (A->format == 1) bytes += double(n) * A->width * (s + 4.0);
  else bytes += double(csr_nnz(ctx, A->rowptr, n)) * (s + 4.0) + double(n + 1) * 4.0;
  return bytes;
}

int check_spmat(size_t n, const trl_hip_spmat* A)
{
  if (!A || (A->format != 0 && A->format != 1) || !valid_spmm_variant(A->variant)) return TRL_HIP_EINVAL;
  if (A->format == 0 && n && !A->rowptr) return TRL_HIP_EINVAL;
  if (A->format == 1 && (A->width < 0 || (A->variant > 1))) return TRL_HIP_EINVAL;
  if (n && (A->format == 0 || A->width > 0) && (!A->col || !A->val)) return TRL_HIP_EINVAL;
  if (n > size_t(INT32_MAX)) return TRL_HIP_EINVAL;
  if (A->plane < 0 || (A->plane > 0 && (A->plane_block < 1 || A->plane_block > A->plane))) return TRL_HIP_EINVAL;
  if ((A->share != 0 && A->share != 1) || A->line < 0 || A->tile_lines < 0) return TRL_HIP_EINVAL;
  return 0;
}

} // namespace

int trl_hip_spmm(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, const void* X, void* Y)
{
  if (!ctx || !valid_b(b) || check_spmat(n, A) || (n && (!X || !Y))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_SPMM, ctx->prof_on ? spmm_bytes(ctx, t, n, b, A, 2) : 0.0, t);
  return launch_spmm(ctx, t, n, b, *A, X, Y, nullptr, false, 0, 0, 0);
}

int trl_hip_spmm_axpby(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, const void* X, void* Y, const void* Z, double alpha,
                       double beta, double gamma)
{
  if (!ctx || !valid_b(b) || check_spmat(n, A) || (n && (!X || !Y)) || X == Y) return TRL_HIP_EINVAL;
  if (!n) return 0;
  // algorithmic panels: X (gathered rows and the beta term read it once), Z, Y
  Prof pr(ctx, TRL_HIP_PROF_SPMM, ctx->prof_on ? spmm_bytes(ctx, t, n, b, A, Z ? 3 : 2) : 0.0, t);
  return launch_spmm(ctx, t, n, b, *A, X, Y, Z, true, alpha, beta, gamma);
}

int trl_hip_spmm_panels(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, int npanels, const void* X, void* Y)
{
  if (!ctx || !valid_b(b) || npanels < 1 || check_spmat(n, A) || (n && (!X || !Y))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_SPMM, ctx->prof_on ? spmm_bytes(ctx, t, n, b, A, 0) + 2.0 * npanels * panel_bytes(t, n, b) : 0.0, t);
  return launch_spmm(ctx, t, n, b, *A, X, Y, nullptr, false, 0, 0, 0, npanels);
}

int trl_hip_spmm_axpby_panels(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, int npanels, const void* X, void* Y, const void* Z,
                              double alpha, double beta, double gamma)
{
  if (!ctx || !valid_b(b) || npanels < 1 || check_spmat(n, A) || (n && (!X || !Y))) return TRL_HIP_EINVAL;
  const size_t es = t == TRL_HIP_F32 ? 4 : 8;
  const char *x = static_cast<const char*>(X), *y = static_cast<const char*>(Y);
  const size_t span = size_t(npanels) * n * b * es;
  if (x < y + span && y < x + span) return TRL_HIP_EINVAL; // X and Y must not overlap
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_SPMM, ctx->prof_on ? spmm_bytes(ctx, t, n, b, A, 0) + (Z ? 3.0 : 2.0) * npanels * panel_bytes(t, n, b) : 0.0, t);
  return launch_spmm(ctx, t, n, b, *A, X, Y, Z, true, alpha, beta, gamma, npanels);
}

int trl_hip_spmm_axpby_forced(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* A, int npanels, const void* X, void* Y, const void* Z,
                              double alpha, double beta, double gamma, const void* F, const double* g)
{
  if (!ctx || !valid_b(b) || npanels < 1 || check_spmat(n, A) || (n && (!X || !Y || !F || !g))) return TRL_HIP_EINVAL;
  const size_t es = t == TRL_HIP_F32 ? 4 : 8;
  const char *x = static_cast<const char*>(X), *y = static_cast<const char*>(Y), *f = static_cast<const char*>(F);
  const size_t span = size_t(npanels) * n * b * es;
  if (x < y + span && y < x + span) return TRL_HIP_EINVAL; // X and Y must not overlap
  if (f < y + span && y < f + span) return TRL_HIP_EINVAL; // nor F and Y
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_SPMM, ctx->prof_on ? spmm_bytes(ctx, t, n, b, A, 0) + (Z ? 4.0 : 3.0) * npanels * panel_bytes(t, n, b) : 0.0, t);
  return launch_spmm(ctx, t, n, b, *A, X, Y, Z, true, alpha, beta, gamma, npanels, F, g);
}

int trl_hip_residual_split(trl_hip_ctx* ctx, size_t n, int b, int npanels, const double* AX, const double* X, const double* theta, double s, float* R,
                           float* E)
{
  if (!ctx || !valid_b(b) || npanels < 1 || npanels > 65535 || (n && (!AX || !X || !theta || !R || !E))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_PANEL, ctx->prof_on ? double(npanels) * (2.0 * panel_bytes(TRL_HIP_F64, n, b) + 2.0 * panel_bytes(TRL_HIP_F32, n, b)) : 0.0,
          TRL_HIP_F64);
  return launch_residual_split(ctx, n, b, npanels, AX, X, theta, s, R, E);
}

int trl_hip_panel_combine(trl_hip_ctx* ctx, size_t n, int b, int npanels, double* X, const double* h, const float* E)
{
  if (!ctx || !valid_b(b) || npanels < 1 || npanels > 65535 || (n && (!X || !h || !E))) return TRL_HIP_EINVAL;
  if (!n) return 0;
  Prof pr(ctx, TRL_HIP_PROF_PANEL, ctx->prof_on ? double(npanels) * (2.0 * panel_bytes(TRL_HIP_F64, n, b) + panel_bytes(TRL_HIP_F32, n, b)) : 0.0,
          TRL_HIP_F64);
  return launch_combine(ctx, n, b, npanels, X, h, E);
}

int trl_hip_spmm_csr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const int32_t* rowptr, const int32_t* col, const void* val, const void* X, void* Y,
                     const void* dscale, int variant)
{
  const trl_hip_spmat A{0, rowptr, 0, col, val, dscale, variant, 0, 0};
  return trl_hip_spmm(ctx, t, n, b, &A, X, Y);
}

int trl_hip_spmm_csr_axpby(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const int32_t* rowptr, const int32_t* col, const void* val, const void* dscale,
                           const void* X, void* Y, const void* Z, double alpha, double beta, double gamma)
{
  const trl_hip_spmat A{0, rowptr, 0, col, val, dscale, 0, 0, 0};
  return trl_hip_spmm_axpby(ctx, t, n, b, &A, X, Y, Z, alpha, beta, gamma);
}

int trl_hip_spmm_ell(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int width, const int32_t* ell_col, const void* ell_val, const void* X, void* Y,
                     const void* dscale)
{
  const trl_hip_spmat A{1, nullptr, width, ell_col, ell_val, dscale, 0, 0, 0};
  return trl_hip_spmm(ctx, t, n, b, &A, X, Y);
}

int trl_hip_csr_max_row(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, int* width)
{
  if (!ctx || !width || (n && !rowptr) || n > size_t(INT32_MAX)) return TRL_HIP_EINVAL;
  TRL_HIP_NO_CAPTURE(ctx);
  *width = 0;
  if (!n) return 0;
  int* d = ctx->flags + 16;
  TRL_HIP_TRY(hipMemsetAsync(d, 0, sizeof(int), ctx->stream));
  int rc = launch_csr_max_row(ctx, n, rowptr, d);
  if (rc) return rc;
  TRL_HIP_TRY(hipMemcpyAsync(width, d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  TRL_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return 0;
}

int trl_hip_csr_to_ell(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, const void* val, int width, int32_t* ell_col,
                       void* ell_val)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx || width < 0 || n > size_t(INT32_MAX) || (n && (!rowptr || (width && (!ell_col || !ell_val))))) return TRL_HIP_EINVAL;
  if (!n || !width) return 0;
  return launch_csr_to_ell(ctx, t, n, rowptr, col, val, width, ell_col, ell_val);
}

} // extern "C"

namespace {

// Autotuner of the operator kernel: times candidate descriptors on scratch
// panels (HIP events around a few back-to-back launches) for the plain
// product, the fused axpby form (Z = Y, the Chebyshev filter's in-place
// step) or that form on npanels panels at once, and keeps the fastest.
// Candidates, pruned so that tuning stays a small part of a solve:
//   - the CSR variants (1, -2, -4, 2, 4, 8), all of them only while the
//     first is within 1.5x of the ELL default;
//   - the ELL variants (1, -2, -4) when an ELL copy is given;
//   - for the fastest of those in index order, blocked row orders (plane =
//     the matrix bandwidth, blocks of about 1/4, 1/8, 1/16 of a plane).
// Several panels: only ELL rows of one chunk share the matrix between
// panels, so the ELL variants (one lane per row only for narrow rows) are
// tried in blocks of 1/8 of a plane (index order when the matrix has no far
// off-diagonals), one timed launch each.
int spmm_tune_impl(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* csr, const trl_hip_spmat* ell, bool axpby, int npanels,
                   trl_hip_spmat* best, double* best_us)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx || !valid_b(b) || !best || npanels < 1 || check_spmat(n, csr) || csr->format != 0 || (ell && (check_spmat(n, ell) || ell->format != 1)))
    return TRL_HIP_EINVAL;
  *best = *csr;
  if (best_us) *best_us = 0.0;
  if (!n) return 0;
  const bool verbose = std::getenv("TRL_HIP_VERBOSE") != nullptr;
  using clock = std::chrono::steady_clock;
  const auto t_start = clock::now();
  const size_t bytes = size_t(npanels) * n * size_t(b) * (t == TRL_HIP_F32 ? 4 : 8);
  void *X = nullptr, *Y = nullptr;
  int rc = trl_hip_malloc(ctx, bytes, &X);
  if (!rc) rc = trl_hip_malloc(ctx, bytes, &Y);
  const double alloc_ms = std::chrono::duration<double, std::milli>(clock::now() - t_start).count();
  const void* Z = axpby ? Y : nullptr; // in place, as the Chebyshev filter calls it
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = TRL_HIP_ENOMEM;
  double best_t = 1e300;
  // Several-panel candidates and small operators: an untimed launch, then 1
  // (several panels) or 3 timed ones.  One-panel candidates on large
  // operators (n b >= 2^24: launches of ~0.1 ms, which differ by more than
  // 10% between candidates): one timed launch, with no untimed one before it
  // once a first candidate has run (the tune runs inside every solve), and 3
  // more if it took under 50 us.
  bool warmed = false;
  auto timed = [&](const trl_hip_spmat& A, int reps) -> double {
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) return -1.0;
    for (int r = 0; r < reps; ++r) (void)launch_spmm(ctx, t, n, b, A, X, Y, Z, axpby, 0.5, -0.25, 0.125, npanels);
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
    return 1e3 * ms / reps;
  };
  auto time_one = [&](const trl_hip_spmat& A) -> double {
    (void)hipGetLastError();
    const bool direct = npanels == 1 && warmed && n * size_t(b) >= (size_t(1) << 24);
    if (!direct) {
      if (launch_spmm(ctx, t, n, b, A, X, Y, Z, axpby, 0.5, -0.25, 0.125, npanels)) {
        (void)hipGetLastError();
        return -1.0; // variant not applicable to this b / dtype
      }
      warmed = true;
      return timed(A, npanels > 1 ? 1 : 3);
    }
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) return -1.0;
    if (launch_spmm(ctx, t, n, b, A, X, Y, Z, axpby, 0.5, -0.25, 0.125, npanels)) {
      (void)hipGetLastError();
      return -1.0; // variant not applicable to this b / dtype
    }
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
    const double us = 1e3 * ms;
    return us >= 50.0 ? us : timed(A, 3);
  };
  auto consider = [&](const trl_hip_spmat& A) -> double {
    const double us = time_one(A);
    if (verbose)
      std::fprintf(stderr, "[tune] %s b=%d %s x%d %s/v%d/plane%d:%d share%d tile%dx%d %.1f us\n", t == TRL_HIP_F32 ? "f32" : "f64", b,
                   axpby ? "axpby" : "apply", npanels, A.format ? "ell" : "csr", A.variant, A.plane, A.plane_block, A.share, A.tile_lines, A.line, us);
    if (us > 0.0 && us < best_t) {
      best_t = us;
      *best = A;
    }
    return us;
  };
  // the matrix bandwidth max |col - row|: the plane of the blocked row orders
  int bw = 0;
  if (!rc) {
    int* d = ctx->flags + 16;
    if (hipMemsetAsync(d, 0, sizeof(int), ctx->stream) != hipSuccess) rc = TRL_HIP_EINVAL;
    if (!rc) rc = launch_csr_bandwidth(ctx, n, csr->rowptr, csr->col, d);
    if (!rc && (hipMemcpyAsync(&bw, d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess))
      rc = TRL_HIP_EINVAL;
  }
  auto with = [&](const trl_hip_spmat& base, int variant, int plane, int block) {
    trl_hip_spmat A = base;
    A.variant = variant;
    A.dscale = csr->dscale;
    A.plane = plane;
    A.plane_block = block;
    return A;
  };
  // a blocked order with about 1/k of a plane per block (whole workgroups), or
  // plane 0 when the far off-diagonals are within a few workgroups
  auto block_of = [&](int variant, int k) {
    // rows per workgroup: the one-panel kernels', or those of the ELL panel kernel
    const int R = spmm_rows_per_block(t, b, variant, npanels > 1);
    if (R <= 0 || bw < 8 * R || size_t(bw) >= n) return 0;
    return std::min(bw, ((bw + k - 1) / k + R - 1) / R * R);
  };
  if (!rc && npanels > 1) {
    // one lane per row only for rows of at most 16 bytes (wider rows hold
    // too many registers across the panel loop)
    const bool narrow = size_t(b) * (t == TRL_HIP_F32 ? 4 : 8) <= 16;
    // index order only where no blocked order applies: the rows of the
    // panels gathered between two uses of a row two planes apart do not
    // fit in L2 (measured on the 160^3 stencil, 7 panels: 1.14x to 1.36x
    // slower than blocks of 1/8 of a plane for every variant)
    if (ell) {
      for (int v : {-2, -4, 1}) {
        if (v == 1 && !narrow) continue;
        const int pb = block_of(v, 8);
        consider(pb ? with(*ell, v, bw, pb) : with(*ell, v, 0, 0));
      }
      // the fastest of those with in-wave gathers shared between lanes, and
      // in tiles of 4 lines (the line length: the largest offset below half
      // the bandwidth) where the tiles cover the plane blocks exactly
      if (best_t < 1e300 && best->format == 1 && ell->width <= 8) {
        int line = 0;
        if (best->plane > 0) {
          int* d = ctx->flags + 16;
          if (hipMemsetAsync(d, 0, sizeof(int), ctx->stream) != hipSuccess || launch_csr_bandwidth_below(ctx, n, csr->rowptr, csr->col, bw / 2, d) ||
              hipMemcpyAsync(&line, d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
            line = 0;
        }
        const trl_hip_spmat base = *best;
        for (int sh : {0, 1})
          for (int ty : {0, 4}) {
            if (sh == 0 && ty == 0) continue;
            if (ty && !line) continue;
            trl_hip_spmat A = base;
            A.share = sh;
            A.line = ty ? line : 0;
            A.tile_lines = ty;
            consider(A);
          }
      }
    }
    else
      consider(with(*csr, 0, 0, 0));
  }
  else if (!rc) {
    double ell_first = -1.0;
    if (ell)
      for (int v : {-2, 1, -4}) {
        const double us = consider(with(*ell, v, 0, 0));
        if (ell_first < 0.0) ell_first = us;
      }
    const double csr_first = consider(with(*csr, 1, 0, 0));
    if (!(ell_first > 0.0 && csr_first > 1.5 * ell_first))
      for (int v : {-2, -4, 2, 4, 8}) consider(with(*csr, v, 0, 0));
    if (best_t < 1e300) {
      const trl_hip_spmat base = *best;
      for (int k : {4, 8, 16}) {
        const int pb = block_of(base.variant, k);
        if (pb) consider(with(base, base.variant, bw, pb));
      }
    }
  }
  if (best_us && best_t < 1e300) *best_us = best_t;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (X) trl_hip_free(ctx, X);
  if (Y) trl_hip_free(ctx, Y);
  if (verbose)
    std::fprintf(stderr, "[tune] %s b=%d %s x%d: %.1f ms (scratch %.1f ms)\n", t == TRL_HIP_F32 ? "f32" : "f64", b, axpby ? "axpby" : "apply", npanels,
                 std::chrono::duration<double, std::milli>(clock::now() - t_start).count(), alloc_ms);
  return rc;
}

} // namespace

extern "C" {

size_t trl_hip_pcg_workspace(size_t n, int b) { return valid_b(b) ? pcg_workspace_bytes(n, b) : 0; }

int trl_hip_pcg(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* K, const void* dinv, const void* rhs, void* X, double tol,
                int maxit, void* work, int* iters, double* relres)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (!ctx || t != TRL_HIP_F64 || !valid_b(b) || check_spmat(n, K) || !(tol > 0.0) || maxit < 0 || (n && (!dinv || !rhs || !X || !work)) || rhs == X)
    return TRL_HIP_EINVAL;
  if (!n) return 0;
  return launch_pcg(ctx, n, b, *K, static_cast<const double*>(dinv), static_cast<const double*>(rhs), static_cast<double*>(X), tol, maxit, work, iters,
                    relres);
}

int trl_hip_spmm_tune(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* csr, const trl_hip_spmat* ell, trl_hip_spmat* best,
                      double* best_us)
{
  return spmm_tune_impl(ctx, t, n, b, csr, ell, false, 1, best, best_us);
}

int trl_hip_spmm_tune_axpby(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const trl_hip_spmat* csr, const trl_hip_spmat* ell, trl_hip_spmat* best,
                            double* best_us)
{
  return spmm_tune_impl(ctx, t, n, b, csr, ell, true, 1, best, best_us);
}

int trl_hip_spmm_tune_panels(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, int npanels, const trl_hip_spmat* csr, const trl_hip_spmat* ell,
                             trl_hip_spmat* best, double* best_us)
{
  return spmm_tune_impl(ctx, t, n, b, csr, ell, true, npanels, best, best_us);
}

int trl_hip_small_copy(trl_hip_ctx* ctx, int b, const double* src, double* dst, int transpose)
{
  if (!ctx || b < 1 || b > 16 || !src || !dst) return TRL_HIP_EINVAL;
  if (!transpose) return trl_hip_memcpy_d2d(ctx, dst, src, size_t(b) * b * sizeof(double));
  Prof pr(ctx, TRL_HIP_PROF_SMALL, 0.0);
  return launch_small(ctx, 1, b, src, nullptr, dst, nullptr);
}

int trl_hip_small_zero(trl_hip_ctx* ctx, int b, double* dst)
{
  if (!ctx || b < 1 || b > 16 || !dst) return TRL_HIP_EINVAL;
  Prof pr(ctx, TRL_HIP_PROF_SMALL, 0.0);
  return launch_small(ctx, 0, b, nullptr, nullptr, dst, nullptr);
}

int trl_hip_small_set_diagonal(trl_hip_ctx* ctx, int b, double* dst, const double* host_values)
{
  if (!ctx || b < 1 || b > 16 || !dst || !host_values) return TRL_HIP_EINVAL;
  Prof pr(ctx, TRL_HIP_PROF_SMALL, 0.0);
  return launch_small(ctx, 2, b, nullptr, nullptr, dst, host_values);
}

int trl_hip_small_mult(trl_hip_ctx* ctx, int b, const double* A, const double* B, double* C)
{
  if (!ctx || b < 1 || b > 16 || !A || !B || !C) return TRL_HIP_EINVAL;
  Prof pr(ctx, TRL_HIP_PROF_SMALL, 0.0);
  return launch_small(ctx, 3, b, A, B, C, nullptr);
}

int trl_hip_residual_sq(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, int b, const void* X, const void* Y, const void* mdiag, const double* lambda, double* out)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx || !valid_b(b) || !out || (n && (!X || !Y))) return TRL_HIP_EINVAL;
  return launch_residual_sq(ctx, t, n, b, X, Y, mdiag, lambda, out);
}

int trl_hip_region_begin(trl_hip_ctx* ctx)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx) return TRL_HIP_EINVAL;
  for (auto& e : ctx->region)
    if (!e) TRL_HIP_TRY(hipEventCreate(&e));
  TRL_HIP_TRY(hipEventRecord(ctx->region[0], ctx->stream));
  return 0;
}

int trl_hip_region_end(trl_hip_ctx* ctx, double* seconds)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx || !seconds || !ctx->region[0]) return TRL_HIP_EINVAL;
  TRL_HIP_TRY(hipEventRecord(ctx->region[1], ctx->stream));
  TRL_HIP_TRY(hipEventSynchronize(ctx->region[1]));
  float ms = 0.f;
  TRL_HIP_TRY(hipEventElapsedTime(&ms, ctx->region[0], ctx->region[1]));
  *seconds = double(ms) * 1e-3;
  return 0;
}

int trl_hip_prof_enable(trl_hip_ctx* ctx, int on)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx) return TRL_HIP_EINVAL;
  int rc = prof_fold(ctx);
  if (rc) return rc;
  if (on) {
    ctx->nnz_cache.clear();
    for (int f = 0; f < TRL_HIP_PROF_FAMILIES; ++f) {
      ctx->prof_count[f] = 0;
      ctx->prof_seconds[f] = 0.0;
      ctx->prof_bytes[f] = 0.0;
      for (int d = 0; d < 2; ++d) {
        ctx->prof_t_count[f][d] = 0;
        ctx->prof_t_seconds[f][d] = 0.0;
        ctx->prof_t_bytes[f][d] = 0.0;
      }
    }
    ctx->reorth_calls = ctx->reorth_second = 0;
  }
  ctx->prof_on = on ? 1 : 0;
  return 0;
}

int trl_hip_prof_read(trl_hip_ctx* ctx, int family, long long* calls, double* seconds, double* bytes)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx || family < 0 || family >= TRL_HIP_PROF_FAMILIES) return TRL_HIP_EINVAL;
  int rc = prof_fold(ctx);
  if (rc) return rc;
  if (calls) *calls = ctx->prof_count[family];
  if (seconds) *seconds = ctx->prof_seconds[family];
  if (bytes) *bytes = ctx->prof_bytes[family];
  return 0;
}

int trl_hip_prof_read_dtype(trl_hip_ctx* ctx, int family, trl_hip_dtype t, long long* calls, double* seconds, double* bytes)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx || family < 0 || family >= TRL_HIP_PROF_FAMILIES || (t != TRL_HIP_F32 && t != TRL_HIP_F64)) return TRL_HIP_EINVAL;
  int rc = prof_fold(ctx);
  if (rc) return rc;
  if (calls) *calls = ctx->prof_t_count[family][t];
  if (seconds) *seconds = ctx->prof_t_seconds[family][t];
  if (bytes) *bytes = ctx->prof_t_bytes[family][t];
  return 0;
}

int trl_hip_prof_reorth(trl_hip_ctx* ctx, long long* calls, long long* second_passes)
{
  if (ctx) TRL_HIP_NO_CAPTURE(ctx);
  if (!ctx) return TRL_HIP_EINVAL;
  if (calls) *calls = ctx->reorth_calls;
  if (second_passes) *second_passes = ctx->reorth_second;
  return 0;
}

} // extern "C"
