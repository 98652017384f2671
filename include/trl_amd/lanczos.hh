#pragma once
// trl_amd::BlockLanczos — block Lanczos with thick restart over any backend
// satisfying trl's Eigenproblem concept (include/trl/concepts.hh:157-188).
//
// This is the build's own restatement of the reference driver
// include/trl/eigensolvers/lanczos.hh:22-312, kept step for step so that it
// produces the same Krylov basis, projected matrix and converged count:
//   extend(k, m)  lanczos.hh:227-287  (three-term recurrence, full
//                                      reorthogonalisation, CholQR)
//   solve()       lanczos.hh:65-199   (Rayleigh-Ritz, thick restart keeping
//                                      nev/b + 1 Ritz blocks, arrow rebuild,
//                                      one quasi-Lanczos step)
// The reference's driver itself also runs unchanged on trl::hip (the
// conformance build compiles it against these backend types); this copy
// exists because the GPU runtime does not carry the reference headers, and
// it adds backend hooks when the EVP provides them:
//   - evp.rotate(V, m, Y, k_r, W): the Ritz rotation as one contraction
//     instead of k_r*m mult_add calls (lanczos.hh:105-114);
//   - the three-term recurrence's "mult into scratch, then -=" pairs as one
//     pass each (subtract_product, subtract_product_transpose);
//   - ritz_vectors(): X = V Y[:, 0:nev], which the reference never forms;
//   - evp.graph_capture_ok(): the restart cycles' block steps captured into a
//     graph once and replayed (extend_cycle), the same kernels in the same
//     order.

#include <concepts>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "trl/hip/context.hh"

namespace trl_amd {

struct EigensolverParams {
  unsigned int nev;
  unsigned int ncv;
  unsigned int max_restarts = 1000;
  double tolerance = 1e-8;
};

struct EigensolverResult {
  bool converged;
  unsigned int iterations;
  unsigned int n_op_apply;
};

// ModifiedGS of reorthogonalization.hh:12-22, for backends without a fused
// multi-block product.
struct ModifiedGS {
  template <typename EVP, typename BMV>
  void operator()(EVP& evp, BMV& V, unsigned int count, typename BMV::BlockView V_next, typename BMV::BlockMatrix::BlockView tmp) const
  {
    for (unsigned int j = 0; j < count; ++j) {
      auto Vj = V.block_view(j);
      evp.dot(Vj, V_next, tmp);
      V_next.subtract_product(Vj, tmp);
    }
  }
};

template <class EVP, class Reorth = ModifiedGS>
class BlockLanczos {
public:
  using BMV = typename EVP::BlockMultivector;
  using Scalar = typename BMV::Scalar;
  static constexpr unsigned int blocksize = BMV::blocksize;

  BlockLanczos(std::shared_ptr<EVP> evp_, const EigensolverParams& params)
      : evp(std::move(evp_)), nev(params.nev), ncv(params.ncv), max_restarts(params.max_restarts),
        V(evp->create_multivector(evp->size(), ncv + blocksize)), W(evp->create_multivector(evp->size(), ncv + blocksize)),
        T(evp->create_blockmatrix(ncv / blocksize, ncv / blocksize)), B(evp->create_blockmatrix(1, 2))
  {
    // the same three checks, in the same order, as lanczos.hh:45-58
    if (ncv + blocksize > evp->size())
      throw std::invalid_argument("ncv (" + std::to_string(ncv) + ") + blocksize (" + std::to_string(blocksize) + ") exceeds problem dimension (" +
                                  std::to_string(evp->size()) + "). Krylov subspace would be exhausted. Reduce ncv to at most " +
                                  std::to_string(evp->size() - blocksize) + ".");
    if (nev >= ncv) throw std::invalid_argument("nev must be strictly smaller than ncv");
    const unsigned int min_blocks = nev / blocksize + 2;
    if (ncv / blocksize < min_blocks)
      throw std::invalid_argument("ncv (" + std::to_string(ncv) + ") is too small for thick restart with nev=" + std::to_string(nev) +
                                  " and blocksize=" + std::to_string(blocksize) + ". Minimum required: ncv >= " + std::to_string(min_blocks * blocksize) +
                                  ".");
    if constexpr (requires { evp->set_tolerance(params.tolerance); }) evp->set_tolerance(params.tolerance);
  }

  EigensolverResult solve()
  {
    EigensolverResult res{false, 0, 0};
    return run(0, res);
  }

  // Continues from a thick-restart state set up by the caller, as solve()
  // leaves it after a restart (lanczos.hh:97-143): V[0:kr] the Ritz blocks,
  // V[kr] the next Lanczos block, and T = diag(theta) on the blocks (i, i),
  // i < kr, plus the arrow T(kr, i) = T(i, kr)^T (kr = nev/b + 1, all other
  // blocks of T[0:kr+1, 0:kr+1] zero).  Runs the quasi-Lanczos step from
  // block kr (lanczos.hh:150-178), then the solve loop.  With it a solve can
  // be warm-started from the restart state of another run (for instance the
  // same problem solved in lower precision).
  EigensolverResult solve_restarted()
  {
    EigensolverResult res{false, 0, 0};
    const unsigned int kr = nev / blocksize + 1;
    if (kr >= ncv / blocksize) throw std::invalid_argument("solve_restarted: no room for a thick restart");
    unsigned int k = kr;
    restarted_ = true;
    if (!quasi_step(k, res)) return res;
    return run(k, res);
  }

  // Grow a block-k Lanczos factorisation to block m; returns the number of
  // operator applications (lanczos.hh:227-287).
  unsigned int extend(unsigned int k, unsigned int m)
  {
    if (k >= m || m > ncv / blocksize) throw std::invalid_argument("extend: need k < m <= ncv/blocksize");
    unsigned int napp = 0;
    auto W0 = W.block_view(0);
    auto beta = B.block_view(0, 0);
    auto Z0 = B.block_view(0, 1);
    if (k == 0) evp->orthonormalize(V.block_view(0), Z0);
    for (unsigned int i = k; i < m; ++i) {
      auto Vc = V.block_view(i);
      auto Vn = V.block_view(i + 1);
      evp->apply(Vc, Vn); // V_{i+1} = A V_i
      ++napp;
      if (i > 0) sub_product_transpose(Vn, V.block_view(i - 1), beta, W0); // V_{i+1} -= V_{i-1} beta^T
      auto Tii = T.block_view(i, i);                                        // T_ii = <V_i, V_{i+1}>
      evp->dot(Vc, Vn, Tii);
      sub_product(Vn, Vc, Tii, W0); // V_{i+1} -= V_i T_ii
      reorth_(*evp, V, i + 1, Vn, Z0);
      evp->orthonormalize(Vn, beta); // V_{i+1} beta = old V_{i+1}
      if (i + 1 < T.block_rows()) {
        T.block_view(i + 1, i).copy_from(beta);
        T.block_view(i, i + 1).copy_from_transpose(beta);
      }
    }
    return napp;
  }

  // True if solve() stopped after a thick restart: the leading nev/b + 1
  // blocks of the basis then already are the Ritz vectors of the last
  // Rayleigh-Ritz step; otherwise they are V Y.
  bool basis_holds_ritz_blocks() const { return restarted_; }

  // Ritz vectors of the last Rayleigh-Ritz step: X[0:nblk] = V[0:m] Y[:, 0:nblk]
  // with nblk = ceil(nev / b) blocks of X.
  void ritz_vectors(BMV& X)
  {
    const unsigned int m = ncv / blocksize;
    const unsigned int nblk = (nev + blocksize - 1) / blocksize;
    if (X.blocks() < nblk) throw std::invalid_argument("ritz_vectors: X too small");
    if (restarted_) {
      for (unsigned int j = 0; j < nblk; ++j) X.block_view(j).copy_from(V.block_view(j));
      return;
    }
    const auto& Y = evp->get_current_eigenvectors();
    if constexpr (requires { evp->rotate(V, m, Y, nblk, X); }) {
      evp->rotate(V, m, Y, nblk, X);
    }
    else {
      for (unsigned int j = 0; j < nblk; ++j) {
        auto Xj = X.block_view(j);
        Xj.set_zero();
        for (unsigned int i = 0; i < m; ++i) V.block_view(i).mult_add(Y.block_view(i, j), Xj);
      }
    }
  }

  auto initial_block() { return V.block_view(0); }
  auto& get_basis() { return V; }
  auto& get_T() { return T; }
  auto& get_B() { return B; }
  EVP& get_evp() { return *evp; }
  // Converged count of the last Rayleigh-Ritz step.  (The reference's
  // get_nconv() always returns 0: lanczos.hh:121,217.)
  unsigned int get_nconv() const { return last_nconv; }
  // Restart cycles run from a captured graph so far.
  unsigned int graph_cycles() const { return graph_cycles_; }

private:
  // The solve loop of lanczos.hh:65-199 from block k: extend to m,
  // Rayleigh-Ritz, and while not converged the thick restart (Ritz rotation,
  // arrow rebuild, quasi-Lanczos step).
  EigensolverResult run(unsigned int k, EigensolverResult res)
  {
    const unsigned int m = ncv / blocksize;
    const unsigned int kr = nev / blocksize + 1;
    auto beta = B.block_view(0, 0);

    while (res.iterations < max_restarts) {
      ++res.iterations;
      res.n_op_apply += k == 0 ? extend(k, m) : extend_cycle(k, m);
      restarted_ = false;

      last_nconv = static_cast<unsigned int>(evp->solve_small_dense(T, beta, nev));
      if (last_nconv >= nev) {
        res.converged = true;
        return res;
      }
      if (kr >= m) break; // no room to restart

      // Ritz rotation: W_j = sum_i V_i Y(i, j), j < kr; W_kr = V_m; swap
      const auto& Y = evp->get_current_eigenvectors();
      if constexpr (requires { evp->rotate(V, m, Y, kr, W); }) {
        evp->rotate(V, m, Y, kr, W);
      }
      else {
        for (unsigned int j = 0; j < kr; ++j) {
          auto Wj = W.block_view(j);
          Wj.set_zero();
          for (unsigned int i = 0; i < m; ++i) V.block_view(i).mult_add(Y.block_view(i, j), Wj);
        }
      }
      W.block_view(kr).copy_from(V.block_view(m));
      std::swap(V, W);
      restarted_ = true;

      // T <- [diag(theta) arrow; arrow^T ...]
      for (unsigned int i = 0; i < kr; ++i)
        for (unsigned int j = 0; j < kr; ++j) {
          auto Tij = T.block_view(i, j);
          Tij.set_zero();
          if (i == j) Tij.set_diagonal(evp->get_eigenvalues_block(i));
        }
      for (unsigned int i = 0; i < kr; ++i) {
        auto Tki = T.block_view(kr, i);
        beta.mult(Y.block_view(m - 1, i), Tki);
        T.block_view(i, kr).copy_from_transpose(Tki);
      }

      k = kr;
      if (!quasi_step(k, res)) break;
    }
    return res;
  }

  // extend(k, m) of a restart cycle.  Every restart cycle runs the same block
  // steps on the same buffers (V alternates between the two the driver swaps),
  // so where the EVP allows it the steps are captured into a graph once per
  // buffer and replayed: one launch per cycle instead of ~20 per block step.
  // The first cycle (k = 0) runs as is and sizes every scratch buffer; a
  // capture that fails (a step needed the host) runs the steps directly and
  // turns capturing off; a graph whose scratch the context or the EVP has
  // since replaced is captured again.
  unsigned int extend_cycle(unsigned int k, unsigned int m)
  {
    if constexpr (requires(EVP& e) {
                    { e.graph_capture_ok() } -> std::convertible_to<bool>;
                    e.graphs();
                    e.scratch_epoch();
                    e.context();
                  }) {
      if (graphs_ok_ && evp->graphs() && evp->graph_capture_ok()) {
        const void* base = static_cast<const void*>(V.data());
        for (auto& cg : cycle_graphs_)
          if (cg.base == base && cg.k == k && cg.m == m && cg.epoch == evp->scratch_epoch()) {
            if (cg.graph.launch(evp->context())) {
              ++graph_cycles_;
              return m - k;
            }
            cg.base = nullptr; // stale: captured again below
          }
        trl::hip::Graph g;
        try {
          g = trl::hip::capture(evp->context(), [&] { extend(k, m); });
        }
        catch (...) {
          g = trl::hip::Graph();
        }
        if (g && g.launch(evp->context())) {
          std::erase_if(cycle_graphs_, [&](const CycleGraph& c) { return c.base == base || c.base == nullptr; });
          cycle_graphs_.push_back(CycleGraph{base, k, m, evp->scratch_epoch(), std::move(g)});
          ++graph_cycles_;
          return m - k;
        }
        graphs_ok_ = false;
      }
    }
    return extend(k, m);
  }

  // One quasi-Lanczos step from block k after a restart (lanczos.hh:150-178):
  // V_{k+1} from A V_k against the arrow and all earlier blocks; k becomes
  // k + 1.  False when T has no room for the step's block.
  bool quasi_step(unsigned int& k, EigensolverResult& res)
  {
    auto beta = B.block_view(0, 0);
    auto Z0 = B.block_view(0, 1);
    auto W0 = W.block_view(0);
    auto Vk = V.block_view(k);
    auto Vk1 = V.block_view(k + 1);
    evp->apply(Vk, Vk1);
    ++res.n_op_apply;
    auto Tkk = T.block_view(k, k);
    evp->dot(Vk, Vk1, Tkk);
    sub_product(Vk1, Vk, Tkk, W0);
    if (k > 0) sub_product_transpose(Vk1, V.block_view(k - 1), T.block_view(k, k - 1), W0);
    reorth_(*evp, V, k + 1, Vk1, Z0);
    evp->orthonormalize(Vk1, beta);
    if (k + 1 < T.block_rows()) {
      T.block_view(k + 1, k).copy_from(beta);
      T.block_view(k, k + 1).copy_from_transpose(beta);
      k += 1;
      return true;
    }
    return false;
  }

  // Y -= X B as mult into scratch + -= (lanczos.hh:162-170, 259-265), or in
  // one pass when the view offers subtract_product (every backend of the
  // concepts does, blockview.hh:189-201)
  template <class View, class MView>
  static void sub_product(View& Y, const View& X, const MView& Bm, View& scratch)
  {
    if constexpr (requires { Y.subtract_product(X, Bm); })
      Y.subtract_product(X, Bm);
    else {
      X.mult(Bm, scratch);
      Y -= scratch;
    }
  }
  // Y -= X B^T (lanczos.hh:166-169, 255-258): one pass when the backend has it
  template <class View, class MView>
  static void sub_product_transpose(View& Y, const View& X, const MView& Bm, View& scratch)
  {
    if constexpr (requires { Y.subtract_product_transpose(X, Bm); })
      Y.subtract_product_transpose(X, Bm);
    else {
      X.mult_transpose(Bm, scratch);
      Y -= scratch;
    }
  }

  std::shared_ptr<EVP> evp;
  unsigned int nev;
  unsigned int ncv;
  unsigned int max_restarts;
  unsigned int last_nconv = 0;
  bool restarted_ = false;
  BMV V;
  BMV W;
  typename BMV::BlockMatrix T;
  typename BMV::BlockMatrix B;
  Reorth reorth_{};
  // captured restart cycles (extend_cycle)
  struct CycleGraph {
    const void* base;
    unsigned int k, m;
    unsigned long long epoch;
    trl::hip::Graph graph;
  };
  std::vector<CycleGraph> cycle_graphs_;
  bool graphs_ok_ = true;
  unsigned int graph_cycles_ = 0;
};

} // namespace trl_amd
