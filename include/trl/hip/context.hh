#pragma once
// trl::hip — C++20 host side of the MI355X backend.  Every operation goes
// through the C-ABI of libtrl_hip.so (trl_hip.h); these headers add the
// types the reference's concepts ask for (include/trl/concepts.hh) and turn
// the C status codes back into the reference's exception types.

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "trl_hip.h"

namespace trl::hip {

// Maps a C-ABI status to the exception the reference raises in the same
// situation: std::invalid_argument (lanczos.hh:46-57, multivector.hh:42),
// std::out_of_range (multivector.hh:87), std::runtime_error
// ("Cholesky factorization failed in orthonormalize", evp_base.hh:55;
// "Eigendecomposition failed", evp_base.hh:96).
inline void throw_on(int rc, const char* what)
{
  if (rc == TRL_HIP_OK) return;
  std::string msg = std::string(what) + ": " + trl_hip_strerror(rc);
  if (rc == TRL_HIP_EINVAL) throw std::invalid_argument(msg);
  if (rc == TRL_HIP_ERANGE) throw std::out_of_range(msg);
  if (rc == TRL_HIP_ECHOL) throw std::runtime_error("Cholesky factorization failed in orthonormalize");
  if (rc == TRL_HIP_EEIG) throw std::runtime_error("Eigendecomposition failed");
  throw std::runtime_error(msg);
}

template <class S>
constexpr trl_hip_dtype dtype_of()
{
  static_assert(std::is_same_v<S, float> || std::is_same_v<S, double>, "panels are float or double");
  return std::is_same_v<S, float> ? TRL_HIP_F32 : TRL_HIP_F64;
}

// Shared handle to one device context (one in-order stream).
class Context {
public:
  explicit Context(int device = 0)
  {
    trl_hip_ctx* c = nullptr;
    throw_on(trl_hip_ctx_create(device, &c), "trl_hip_ctx_create");
    h_ = std::shared_ptr<trl_hip_ctx>(c, [](trl_hip_ctx* p) { trl_hip_ctx_destroy(p); });
  }
  // Wraps an existing context without owning it.
  static Context borrow(trl_hip_ctx* c)
  {
    Context x(nullptr);
    x.h_ = std::shared_ptr<trl_hip_ctx>(c, [](trl_hip_ctx*) {});
    return x;
  }

  trl_hip_ctx* get() const { return h_.get(); }

  // Waits for queued work and raises a pending Cholesky breakdown.
  void sync() const
  {
    int piv = 0;
    throw_on(trl_hip_ctx_check(h_.get(), &piv), "sync");
  }

private:
  explicit Context(std::nullptr_t) {}
  std::shared_ptr<trl_hip_ctx> h_;
};

// A captured graph of queued work (trl_hip_graph_*), owned.
class Graph {
public:
  Graph() = default;
  explicit Graph(trl_hip_graph* g) : g_(g, [](trl_hip_graph* p) { trl_hip_graph_destroy(p); }) {}
  explicit operator bool() const { return bool(g_); }
  // Queues one replay; false when the graph is stale (the context replaced a
  // scratch buffer it names), so that the caller captures it again.
  bool launch(const Context& ctx) const
  {
    const int rc = trl_hip_graph_launch(ctx.get(), g_.get());
    if (rc == TRL_HIP_ESTALE) return false;
    throw_on(rc, "graph launch");
    return true;
  }

private:
  std::shared_ptr<trl_hip_graph> g_;
};

// Captures the work `fn` queues on the context stream into a graph, without
// running it.  Returns an empty Graph when the capture failed (some call in
// fn needed the host): the work has then not run, and the caller runs fn
// itself.  An exception thrown by fn ends the capture and propagates.
template <class F>
Graph capture(const Context& ctx, F&& fn)
{
  if (trl_hip_graph_begin(ctx.get()) != TRL_HIP_OK) return Graph();
  try {
    fn();
  }
  catch (...) {
    trl_hip_graph* g = nullptr;
    (void)trl_hip_graph_end(ctx.get(), &g);
    trl_hip_graph_destroy(g);
    throw;
  }
  trl_hip_graph* g = nullptr;
  if (trl_hip_graph_end(ctx.get(), &g) != TRL_HIP_OK) return Graph();
  return Graph(g);
}

// RAII device buffer.
template <class T>
class DeviceBuffer {
public:
  DeviceBuffer() = default;
  DeviceBuffer(const Context& ctx, std::size_t count) : ctx_(ctx), count_(count)
  {
    void* p = nullptr;
    throw_on(trl_hip_malloc(ctx_.get(), count * sizeof(T), &p), "trl_hip_malloc");
    p_ = static_cast<T*>(p);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ctx_(o.ctx_), p_(o.p_), count_(o.count_)
  {
    o.p_ = nullptr;
    o.count_ = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
  {
    if (this != &o) {
      release();
      ctx_ = o.ctx_;
      p_ = o.p_;
      count_ = o.count_;
      o.p_ = nullptr;
      o.count_ = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }

  T* data() const { return p_; }
  std::size_t size() const { return count_; }
  const Context& context() const { return ctx_; }

private:
  void release()
  {
    if (p_) trl_hip_free(ctx_.get(), p_);
    p_ = nullptr;
  }
  Context ctx_ = Context::borrow(nullptr);
  T* p_ = nullptr;
  std::size_t count_ = 0;
};

} // namespace trl::hip
