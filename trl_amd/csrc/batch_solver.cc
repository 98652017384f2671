// Batched thick-restart block Lanczos driver (trl_hip_batch_solve_csr).
//
// The restart logic of include/trl/eigensolvers/lanczos.hh:65-199, run in
// lock-step over many problems: every cycle extends all active problems with
// one replayed hipGraph (libtrl_hip.so, trl_hip_batch_extend), runs the
// Rayleigh-Ritz step of each problem on the host (in parallel threads, the
// same trl_amd::rayleigh_ritz as the single-problem path), rotates all
// active problems at once and retires the converged ones.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include <omp.h>

#include "trl_amd/rayleigh_ritz.hh"
#include "trl_hip_solver.h"

namespace {

double now()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Problem {
  std::vector<double> T; // m x m blocks of b x b
  trl_amd::RitzResult rr;
  std::vector<double> dense;
  std::vector<double> arrow;
  bool active = true;
  bool converged = false;
  int error = 0;
  unsigned cycles = 0;
  unsigned napp = 0;
  double t_done = 0.0;
};

template <class F>
void parallel_for(const std::vector<int>& items, unsigned nthreads, F&& f)
{
  if (items.empty()) return;
  nthreads = std::max(1u, std::min<unsigned>(nthreads, unsigned(items.size())));
  if (nthreads == 1) {
    for (int p : items) f(p);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthreads; ++t)
    pool.emplace_back([&]() {
      // the dense eigensolver threads its larger reductions with OpenMP;
      // inside this pool each worker runs them serially (the same results:
      // they do not depend on the thread count) instead of each starting a
      // team of its own
      omp_set_num_threads(1);
      for (size_t i; (i = next.fetch_add(1)) < items.size();) f(items[i]);
    });
  for (auto& th : pool) th.join();
}

struct BatchGuard {
  trl_hip_batch* b = nullptr;
  ~BatchGuard()
  {
    if (b) trl_hip_batch_destroy(b);
  }
};

} // namespace

extern "C" int trl_hip_batch_solve_csr(trl_hip_ctx* ctx, int nprob, const size_t* n, const int32_t* const* rowptr, const int32_t* const* col,
                                       const double* const* val, const double* const* V0, const trl_hip_batch_params* prm, double* evals, double* resid,
                                       double* const* X, trl_hip_solver_result* res, trl_hip_batch_stats* stats)
{
  if (!ctx || nprob < 1 || !n || !rowptr || !col || !val || !V0 || !prm || !evals || !res) return TRL_HIP_EINVAL;
  const int b = prm->b;
  const unsigned nev = prm->nev, ncv = prm->ncv;
  if (b < 1 || ncv % unsigned(b) || prm->order < 0 || prm->order > 2 || prm->reorth < 0 || prm->reorth > 2) return TRL_HIP_EINVAL;
  const unsigned m = ncv / unsigned(b), kr = nev / unsigned(b) + 1;
  // the checks of lanczos.hh:45-58, for every problem
  if (nev >= ncv || m < nev / unsigned(b) + 2) return TRL_HIP_EINVAL;
  for (int p = 0; p < nprob; ++p)
    if (ncv + unsigned(b) > n[p] || !V0[p]) return TRL_HIP_EINVAL;
  const int BB = b * b;
  const double t_setup = now();
  BatchGuard g;
  int rc = trl_hip_batch_create(ctx, nprob, n, rowptr, col, val, b, int(ncv), &g.b);
  if (rc) return rc;
  if ((rc = trl_hip_batch_set_options(g.b, prm->cholqr_passes > 0 ? prm->cholqr_passes : 1, prm->use_graph))) return rc;
  if ((rc = trl_hip_batch_set_reorth(g.b, prm->reorth == 2 ? 1 : 0))) return rc;
  for (int p = 0; p < nprob; ++p)
    if ((rc = trl_hip_batch_set_block(g.b, p, 0, V0[p]))) return rc;
  double tuned_us = 0.0;
  if (prm->tune && (rc = trl_hip_batch_tune(g.b, &tuned_us))) return rc;

  std::vector<Problem> P(nprob);
  for (auto& q : P) {
    q.T.assign(size_t(m) * m * BB, 0.0);
    q.arrow.assign(BB, 0.0);
  }
  std::vector<double> alpha(size_t(nprob) * m * BB), beta(size_t(nprob) * m * BB), arrow(size_t(nprob) * BB, 0.0);
  std::vector<double> Y(size_t(nprob) * m * b * kr * b, 0.0);
  std::vector<int> pflag(nprob), active(nprob, 1);
  const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  unsigned cycles = 0, launches = 0;
  double host_s = 0.0, extend_s = 0.0, rotate_s = 0.0;
  unsigned k = 0;
  int quasi = 0;
  const unsigned nblk = (nev + unsigned(b) - 1) / unsigned(b);
  size_t rows_tiled = 0;
  for (int p = 0; p < nprob; ++p) rows_tiled += n[p];
  unsigned retiles = 0;
  const bool verbose = std::getenv("TRL_HIP_VERBOSE") != nullptr;
  const double t0 = now();
  const double setup_s = t0 - t_setup;
  auto Tblk = [&](Problem& q, unsigned i, unsigned j) { return q.T.data() + (size_t(i) * m + j) * BB; };

  while (cycles < prm->max_restarts) {
    std::vector<int> live;
    for (int p = 0; p < nprob; ++p)
      if (P[p].active) live.push_back(p);
    if (live.empty()) break;
    ++cycles;
    const double te = now();
    if ((rc = trl_hip_batch_extend(g.b, int(k), int(m), quasi, quasi ? arrow.data() : nullptr, alpha.data(), beta.data(), pflag.data()))) return rc;
    extend_s += now() - te;
    launches += 1;
    const double th = now();
    parallel_for(live, nthreads, [&](int p) {
      Problem& q = P[p];
      if (pflag[p]) {
        q.error = TRL_HIP_ECHOL;
        return;
      }
      for (unsigned i = k; i < m; ++i) {
        const double* a = alpha.data() + (size_t(p) * m + i) * BB;
        const double* bb = beta.data() + (size_t(p) * m + i) * BB;
        std::memcpy(Tblk(q, i, i), a, BB * sizeof(double));
        if (i + 1 < m)
          for (int r = 0; r < b; ++r)
            for (int c = 0; c < b; ++c) {
              Tblk(q, i + 1, i)[r * b + c] = bb[r * b + c];
              Tblk(q, i, i + 1)[c * b + r] = bb[r * b + c];
            }
      }
      q.napp += m - k;
      q.cycles = cycles;
      const double* blast = beta.data() + (size_t(p) * m + (m - 1)) * BB;
      try {
        trl_amd::rayleigh_ritz(unsigned(b), m, m, q.T.data(), blast, nev, prm->order, prm->tol, q.rr, q.dense, size_t(kr) * b);
      }
      catch (const std::exception&) {
        q.error = TRL_HIP_EEIG;
        return;
      }
      if (q.rr.nconv >= nev) q.converged = true;
      // rotation coefficients: Y_p(i*b + bi, j*b + bj) = Y(i, j)(bi, bj), j < kr
      double* yp = Y.data() + size_t(p) * m * b * kr * b;
      for (unsigned i = 0; i < m; ++i)
        for (unsigned j = 0; j < kr; ++j)
          for (int bi = 0; bi < b; ++bi)
            for (int bj = 0; bj < b; ++bj) yp[(size_t(i) * b + bi) * kr * b + j * b + bj] = q.rr.Y[((size_t(i) * m + j) * b + bi) * b + bj];
    });
    host_s += now() - th;
    if (verbose) {
      size_t rows_live = 0;
      for (int p : live) rows_live += n[p];
      std::fprintf(stderr, "[batch] cycle %u: %zu problems, %zu rows, extend %.3f ms, host %.3f ms\n", cycles, live.size(), rows_live, 1e3 * (th - te),
                   1e3 * (now() - th));
    }
    bool restart_possible = kr < m && kr + 1 < m;
    for (int p : live)
      if (P[p].error) {
        P[p].active = false;
        active[p] = 0;
      }
    // every problem still in the batch this cycle is rotated: the converged
    // ones get their Ritz vectors in blocks 0..kr-1, the others restart
    if ((rc = trl_hip_batch_set_active(g.b, active.data()))) return rc;
    const double tr = now();
    if ((rc = trl_hip_batch_rotate(g.b, Y.data(), int(kr)))) return rc;
    rotate_s += now() - tr;
    const double tdone = now();
    for (int p : live) {
      Problem& q = P[p];
      if (q.error) {
        if ((rc = trl_hip_batch_save_blocks(g.b, p, int(nblk)))) return rc;
        continue;
      }
      if (q.converged || !restart_possible || cycles >= prm->max_restarts) {
        q.active = false;
        active[p] = 0;
        q.t_done = tdone - t0;
        // its Ritz vectors leave the basis: later cycles only cover active problems
        if ((rc = trl_hip_batch_save_blocks(g.b, p, int(nblk)))) return rc;
        continue;
      }
      // T <- [diag(theta) arrow; arrow^T ...] (lanczos.hh:123-143)
      std::fill(q.T.begin(), q.T.end(), 0.0);
      for (unsigned i = 0; i < kr; ++i)
        for (int c = 0; c < b; ++c) Tblk(q, i, i)[c * b + c] = q.rr.theta[i * b + c];
      const double* blast = beta.data() + (size_t(p) * m + (m - 1)) * BB;
      for (unsigned i = 0; i < kr; ++i) {
        const double* ym = q.rr.Y.data() + ((size_t(m - 1) * m + i) * b) * b; // block (m-1, i)
        double* tki = Tblk(q, kr, i);
        for (int r = 0; r < b; ++r)
          for (int c = 0; c < b; ++c) {
            double s = 0.0;
            for (int x = 0; x < b; ++x) s += blast[r * b + x] * ym[x * b + c];
            tki[r * b + c] = s;
          }
        for (int r = 0; r < b; ++r)
          for (int c = 0; c < b; ++c) Tblk(q, i, kr)[c * b + r] = tki[r * b + c];
      }
      std::memcpy(arrow.data() + size_t(p) * BB, Tblk(q, kr, kr - 1), BB * sizeof(double));
    }
    if ((rc = trl_hip_batch_set_active(g.b, active.data()))) return rc;
    // fewer rows left: new tiles over the active problems only (re-tuned:
    // with a few large problems left, smaller tiles give more workgroups)
    size_t rows_active = 0;
    for (int p = 0; p < nprob; ++p)
      if (P[p].active) rows_active += n[p];
    if (rows_active > 0 && 2 * rows_active < rows_tiled) {
      if (prm->tune)
        rc = trl_hip_batch_tune(g.b, nullptr);
      else {
        int fmt = 0, rpt = 0;
        trl_hip_batch_info(g.b, nullptr, &fmt, &rpt, nullptr);
        rc = trl_hip_batch_set_variant(g.b, fmt, rpt);
      }
      if (rc) return rc;
      rows_tiled = rows_active;
      ++retiles;
    }
    k = kr;
    quasi = 1;
  }
  const double total = now() - t0;

  // results: Ritz values, vectors (blocks 0..nblk-1) and true residuals
  std::vector<double> lam(size_t(nprob) * nblk * b, 0.0), rr(size_t(nprob) * nblk * b, 0.0);
  int first_error = 0;
  for (int p = 0; p < nprob; ++p) {
    const Problem& q = P[p];
    res[p] = trl_hip_solver_result{};
    res[p].converged = q.converged ? 1 : 0;
    res[p].iterations = q.cycles;
    res[p].n_op_apply = q.napp;
    res[p].seconds = q.t_done > 0.0 ? q.t_done : total;
    for (unsigned i = 0; i < nev; ++i) {
      const double v = i < q.rr.theta.size() ? q.rr.theta[i] : 0.0;
      evals[size_t(p) * nev + i] = v;
      lam[size_t(p) * nblk * b + i] = v;
    }
    if (q.error && !first_error) first_error = q.error;
  }
  if (resid) {
    if ((rc = trl_hip_batch_residuals(g.b, int(nblk), lam.data(), rr.data()))) return rc;
    for (int p = 0; p < nprob; ++p)
      for (unsigned i = 0; i < nev; ++i) resid[size_t(p) * nev + i] = rr[size_t(p) * nblk * b + i];
  }
  if (X)
    for (int p = 0; p < nprob; ++p)
      if (X[p])
        for (unsigned j = 0; j < nblk; ++j)
          if ((rc = trl_hip_batch_get_block(g.b, p, int(j), X[p] + size_t(j) * n[p] * b))) return rc;
  if (stats) {
    *stats = trl_hip_batch_stats{};
    stats->seconds = total;
    stats->host_seconds = host_s;
    stats->setup_seconds = setup_s;
    stats->cycles = cycles;
    stats->graph_launches = launches;
    stats->extend_seconds = extend_s;
    stats->retiles = retiles;
    stats->rotate_seconds = rotate_s;
    trl_hip_batch_info(g.b, &stats->ntiles, &stats->format, &stats->rows_per_thread, &stats->tuned_us);
  }
  return first_error;
}
