// libtrl_hip_solver.so: whole-solve C-ABI (include/trl_hip_solver.h) over the
// trl::hip backend types and the trl_amd::BlockLanczos driver.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "trl/hip/csrevp.hh"
#include "trl_amd/dense.hh"
#include "trl_amd/lanczos.hh"
#include "trl_hip_solver.h"

namespace {

using trl::hip::Context;
using trl::hip::throw_on;

int code_of(const std::exception& e)
{
  if (dynamic_cast<const std::invalid_argument*>(&e)) return TRL_HIP_EINVAL;
  if (dynamic_cast<const std::out_of_range*>(&e)) return TRL_HIP_ERANGE;
  const std::string w = e.what();
  if (w.find("Cholesky") != std::string::npos) return TRL_HIP_ECHOL;
  if (w.find("Eigendecomposition") != std::string::npos) return TRL_HIP_EEIG;
  if (w.find("allocation") != std::string::npos) return TRL_HIP_ENOMEM;
  return TRL_HIP_EINVAL;
}

double now()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// True residuals ||K x - lambda M x|| / ||K x|| of nev Ritz vectors stored as
// ceil(nev/B) panels at X (device), computed with the undivided operator K.
template <class S, unsigned B>
void true_residuals(const Context& c, std::size_t n, const int32_t* rowptr, const int32_t* col, const S* val, const S* mass, const S* X, unsigned nev,
                    const double* evals, double* out, int lanes)
{
  constexpr trl_hip_dtype t = trl::hip::dtype_of<S>();
  trl::hip::DeviceBuffer<S> KX(c, n * B);
  const unsigned nblk = (nev + B - 1) / B;
  for (unsigned j = 0; j < nblk; ++j) {
    const S* Xj = X + std::size_t(j) * n * B;
    throw_on(trl_hip_spmm_csr(c.get(), t, n, B, rowptr, col, val, Xj, KX.data(), nullptr, lanes), "residual spmm");
    double lam[16] = {0.0};
    for (unsigned q = 0; q < B && j * B + q < nev; ++q) lam[q] = evals[j * B + q];
    double sums[32];
    throw_on(trl_hip_residual_sq(c.get(), t, n, B, Xj, KX.data(), mass, lam, sums), "residual_sq");
    for (unsigned q = 0; q < B && j * B + q < nev; ++q) out[j * B + q] = std::sqrt(sums[q]) / std::max(std::sqrt(sums[B + q]), 1e-300);
  }
}

struct SolveOut {
  trl_hip_solver_result res{};
};

// Runs one solve with panel type S and block size B.  If X_out is given it
// receives the first nblk_out Ritz blocks.
template <class S, unsigned B>
void solve_typed(const Context& c, std::size_t n, const int32_t* rowptr, const int32_t* col, const S* val, const S* mass, const S* minv,
                 const trl_hip_solver_params& p, const S* V0, double* evals, S* X_out, unsigned nblk_out, double* resid, trl_hip_solver_result& res)
{
  using EVP = trl::hip::CSREVP<S, B>;
  auto evp = std::make_shared<EVP>(c, n, rowptr, col, val, mass, minv);
  evp->set_order(static_cast<trl::hip::EigenvalueOrder>(p.order));
  evp->set_lanes_per_row(p.lanes_per_row);
  if (p.cholqr_passes > 0) evp->set_cholqr_passes(p.cholqr_passes);
  trl_amd::EigensolverParams ep{p.nev, p.ncv, p.max_restarts, p.tol};

  auto run = [&](auto& lz) {
    auto V0b = lz.initial_block();
    throw_on(trl_hip_memcpy_d2d(c.get(), V0b.data_, V0, n * B * sizeof(S)), "copy V0");
    c.sync();
    const double t0 = now();
    const auto r = lz.solve();
    c.sync();
    res.seconds = now() - t0;
    res.converged = r.converged ? 1 : 0;
    res.iterations = r.iterations;
    res.n_op_apply = r.n_op_apply;
    res.host_seconds = evp->host_seconds();
    const auto& ev = evp->get_current_eigenvalues();
    for (unsigned i = 0; i < p.nev && i < ev.size(); ++i) evals[i] = ev[i];
    if (X_out) {
      const unsigned kr = p.nev / B + 1;
      if (lz.basis_holds_ritz_blocks() && nblk_out <= kr) {
        throw_on(trl_hip_memcpy_d2d(c.get(), X_out, lz.get_basis().data(), std::size_t(nblk_out) * n * B * sizeof(S)), "ritz vectors");
      }
      else {
        const auto& Y = evp->get_current_eigenvectors();
        throw_on(trl_hip_rotate(c.get(), EVP::dtype, n, B, int(p.ncv / B), lz.get_basis().data(), Y.data(), int(Y.block_cols()), int(nblk_out), X_out),
                 "ritz vectors");
      }
      if (resid) true_residuals<S, B>(c, n, rowptr, col, val, mass, X_out, p.nev, evals, resid, p.lanes_per_row);
      c.sync();
    }
  };
  if (p.reorth == 1) {
    trl_amd::BlockLanczos<EVP, trl::hip::BlockCGS2> lz(evp, ep);
    run(lz);
  }
  else {
    trl_amd::BlockLanczos<EVP, trl_amd::ModifiedGS> lz(evp, ep);
    run(lz);
  }
}

template <class F>
int dispatch_b(int b, F&& f)
{
  switch (b) {
  case 1:
    f(std::integral_constant<unsigned, 1>{});
    return 0;
  case 2:
    f(std::integral_constant<unsigned, 2>{});
    return 0;
  case 4:
    f(std::integral_constant<unsigned, 4>{});
    return 0;
  case 6:
    f(std::integral_constant<unsigned, 6>{});
    return 0;
  case 8:
    f(std::integral_constant<unsigned, 8>{});
    return 0;
  case 16:
    f(std::integral_constant<unsigned, 16>{});
    return 0;
  default:
    return TRL_HIP_EINVAL;
  }
}

} // namespace

extern "C" {

int trl_hip_dense_syev(int n, const double* A, double* w, double* Z)
{
  if (n < 0 || (n > 0 && (!A || !w || !Z))) return TRL_HIP_EINVAL;
  return trl_amd::dense::syev(n, A, w, Z) == 0 ? 0 : TRL_HIP_EEIG;
}

int trl_hip_solve_csr(trl_hip_ctx* ctx, trl_hip_dtype t, size_t n, const int32_t* rowptr, const int32_t* col, const void* val, const void* mass,
                      const void* mass_inv, const trl_hip_solver_params* p, const void* V0, double* evals, void* X, double* resid,
                      trl_hip_solver_result* res)
{
  if (!ctx || !p || !rowptr || !col || !val || !V0 || !evals || !res) return TRL_HIP_EINVAL;
  if ((mass == nullptr) != (mass_inv == nullptr)) return TRL_HIP_EINVAL;
  if (p->order < 0 || p->order > 2 || p->reorth < 0 || p->reorth > 1) return TRL_HIP_EINVAL;
  *res = trl_hip_solver_result{};
  try {
    const Context c = Context::borrow(ctx);
    const unsigned nblk = (p->nev + unsigned(p->b) - 1) / unsigned(std::max(p->b, 1));
    int rc = 0;
    if (t == TRL_HIP_F64) {
      rc = dispatch_b(p->b, [&](auto Bc) {
        constexpr unsigned B = decltype(Bc)::value;
        solve_typed<double, B>(c, n, rowptr, col, static_cast<const double*>(val), static_cast<const double*>(mass), static_cast<const double*>(mass_inv),
                               *p, static_cast<const double*>(V0), evals, static_cast<double*>(X), nblk, resid, *res);
      });
    }
    else if (t == TRL_HIP_F32) {
      rc = dispatch_b(p->b, [&](auto Bc) {
        constexpr unsigned B = decltype(Bc)::value;
        solve_typed<float, B>(c, n, rowptr, col, static_cast<const float*>(val), static_cast<const float*>(mass), static_cast<const float*>(mass_inv), *p,
                              static_cast<const float*>(V0), evals, static_cast<float*>(X), nblk, resid, *res);
      });
    }
    else
      return TRL_HIP_EINVAL;
    return rc;
  }
  catch (const std::exception& e) {
    return code_of(e);
  }
}

int trl_hip_solve_csr_mixed(trl_hip_ctx* ctx, size_t n, const int32_t* rowptr, const int32_t* col, const double* val64, const float* val32,
                            const double* mass64, const double* minv64, const float* mass32, const float* minv32, const trl_hip_solver_params* p,
                            const trl_hip_mixed_params* mp, const double* V0, double* evals, double* X, double* resid, trl_hip_solver_result* res,
                            trl_hip_solver_result* res32)
{
  if (!ctx || !p || !mp || !rowptr || !col || !val64 || !val32 || !V0 || !evals || !res) return TRL_HIP_EINVAL;
  if ((mass64 == nullptr) != (minv64 == nullptr) || (mass32 == nullptr) != (minv32 == nullptr) || (mass64 == nullptr) != (mass32 == nullptr))
    return TRL_HIP_EINVAL;
  *res = trl_hip_solver_result{};
  trl_hip_solver_result r32{};
  try {
    const Context c = Context::borrow(ctx);
    const int rc = dispatch_b(p->b, [&](auto Bc) {
      constexpr unsigned B = decltype(Bc)::value;
      const unsigned kr = p->nev / B + 1;
      c.sync();
      const double t0 = now();
      // ---- FP32 phase -------------------------------------------------------
      trl::hip::DeviceBuffer<float> V0f(c, n * B);
      throw_on(trl_hip_convert(c.get(), TRL_HIP_F64, V0, TRL_HIP_F32, V0f.data(), n * B), "convert V0");
      trl::hip::DeviceBuffer<float> X32(c, std::size_t(kr) * n * B);
      std::vector<double> ev32(std::max(p->nev, kr * B), 0.0);
      trl_hip_solver_params p32 = *p;
      p32.tol = mp->tol_fp32;
      p32.max_restarts = mp->max_restarts_fp32;
      if (p32.cholqr_passes == 0) p32.cholqr_passes = 2;
      solve_typed<float, B>(c, n, rowptr, col, val32, mass32, minv32, p32, V0f.data(), ev32.data(), X32.data(), kr, nullptr, r32);
      // ---- warm start: V0_64 = X32 G ---------------------------------------
      trl::hip::DeviceBuffer<double> X64(c, std::size_t(kr) * n * B);
      throw_on(trl_hip_convert(c.get(), TRL_HIP_F32, X32.data(), TRL_HIP_F64, X64.data(), std::size_t(kr) * n * B), "convert X32");
      std::mt19937_64 g(mp->seed);
      std::normal_distribution<double> nd;
      std::vector<double> G(std::size_t(kr) * B * B);
      for (auto& x : G) x = nd(g);
      trl::hip::DeviceBuffer<double> Gd(c, G.size());
      throw_on(trl_hip_memcpy_h2d(c.get(), Gd.data(), G.data(), G.size() * sizeof(double)), "upload G");
      trl::hip::DeviceBuffer<double> V064(c, n * B);
      throw_on(trl_hip_rotate(c.get(), TRL_HIP_F64, n, B, int(kr), X64.data(), Gd.data(), 1, 1, V064.data()), "mix Ritz vectors");
      // ---- FP64 phase --------------------------------------------------------
      const unsigned nblk = (p->nev + B - 1) / B;
      solve_typed<double, B>(c, n, rowptr, col, val64, mass64, minv64, *p, V064.data(), evals, X, nblk, resid, *res);
      c.sync();
      res->seconds = now() - t0;
    });
    if (res32) *res32 = r32;
    return rc;
  }
  catch (const std::exception& e) {
    if (res32) *res32 = r32;
    return code_of(e);
  }
}

} // extern "C"
