"""Whole-solve parity of the HIP backend (libtrl_hip_solver.so) against the
CPU oracle and the closed forms, on the same seeded inputs.

Bar (from the benchmark definition): eigenvalues within 1e-10 relative of
the oracle's FP64 run, true residuals ||Ax - lambda M x|| / ||Ax|| at the
solver tolerance, principal angles between the HIP and oracle invariant
subspaces reported and bounded.
"""

import numpy as np
import pytest
import scipy.linalg as sla

import oracle
from trl_amd import _lib as L
from trl_amd import problems as P
from trl_amd.solver import DeviceCSR, EigensolverParams, solve, solve_mixed

pytestmark = pytest.mark.gpu


def v0(ctx, n, b, dt=np.float64, seed=42):
    return ctx.to_device(P.initial_block(n, b, seed).astype(dt))


def max_angle(Xa, Xb):
    Qa, _ = np.linalg.qr(Xa)
    Qb, _ = np.linalg.qr(Xb)
    return float(np.max(sla.subspace_angles(Qa, Qb)))


@pytest.mark.parametrize("reorth", ["mgs", "cgs2", "dgks"])
@pytest.mark.parametrize("b", [1, 2, 4, 8])
@pytest.mark.parametrize("kind", ["diagonal", "laplace1d"])
def test_reference_convergence(ctx, kind, b, reorth):
    # tests/openmp/test_lanczos_convergence.cc: N = 128, nev = 16, ncv = 64,
    # largest |lambda| first, relative 1e-8 against the exact values
    print(f"convergence {kind} b={b} {reorth}", flush=True)
    N, nev = 128, 16
    A = P.diagonal(N) if kind == "diagonal" else P.laplacian_1d(N)
    exact = (N - np.arange(nev)) ** 2.0 if kind == "diagonal" else 2 - 2 * np.cos((N - np.arange(nev)) * np.pi / (N + 1))
    dA = DeviceCSR(ctx, A, precisions=("f64",))
    out = solve(dA, b, EigensolverParams(nev, 4 * nev), v0(ctx, N, b), order="largest_magnitude", reorth=reorth)
    assert out.result["converged"] == 1
    assert np.all(np.abs(out.evals - exact) <= 1e-8 * np.abs(exact))
    ref = oracle.lanczos_solve(A, b, nev, 4 * nev, P.initial_block(N, b), order=0)
    assert np.all(np.abs(out.evals - ref["evals"]) <= 1e-10 * np.abs(ref["evals"]))


@pytest.mark.parametrize("reorth", ["mgs", "cgs2", "dgks"])
def test_laplace2d_128_matches_oracle(ctx, reorth):
    # 2D 5-point Laplacian 128^2, b = 4, 8 smallest, ncv = 256, tol 1e-8
    c = P.config_c1()
    print(f"laplace2d_128 {reorth}", flush=True)
    dA = DeviceCSR(ctx, c.csr, precisions=("f64",))
    out = solve(dA, c.b, EigensolverParams(c.nev, c.ncv, tolerance=c.tol), v0(ctx, c.csr.n, c.b), order="ascending", reorth=reorth)
    assert out.result["converged"] == 1
    ref = oracle.lanczos_solve(c.csr, c.b, c.nev, c.ncv, P.initial_block(c.csr.n, c.b), tol=c.tol, order=1)
    assert ref["converged"]
    rel_or = np.abs(out.evals - ref["evals"]) / np.abs(ref["evals"])
    rel_ex = np.abs(out.evals - c.exact) / c.exact
    print("rel vs oracle", rel_or.max(), "rel vs closed form", rel_ex.max(), "restarts", out.result["iterations"], ref["iterations"])
    assert rel_or.max() <= 1e-10
    assert rel_ex.max() <= 1e-10
    # the Ritz estimate ||beta y_last|| / |theta| < tol that stops the solve
    # (evp_base.hh:125-151) equals ||A x - theta x|| / ||A x|| up to rounding
    # for a standard problem: the true residuals meet the tolerance
    print("true residuals", np.array2string(out.resid, precision=2))
    assert out.resid.max() <= c.tol
    X = out.X.download()  # (nblk, n, b)
    Xh = np.concatenate(list(X), axis=1)[:, : c.nev]
    ang = max_angle(Xh, ref["X"])
    print("max principal angle", ang)
    assert ang < 1e-6


@pytest.mark.parametrize("order", ["ascending", "largest_magnitude"])
def test_structured_rayleigh_ritz_matches_dense(ctx, order):
    # the same 128^2 solve (nt = 256) with the structured Rayleigh-Ritz
    # eigensolver (arrow + band, the default) and with the dense Householder
    # one in every step: every step of the default run takes the structured
    # path, the dense run none; eigenvalues agree to 1e-12 relative, both meet
    # the tolerance, and the host time per step is printed
    c = P.config_c1()
    print(f"structured RR laplace2d_128 order={order}", flush=True)
    dA = DeviceCSR(ctx, c.csr, precisions=("f64",))
    prm = EigensolverParams(c.nev, c.ncv, tolerance=c.tol)
    # one untimed solve first: the first one pays for starting the host
    # thread pool and first touches
    solve(dA, c.b, prm, v0(ctx, c.csr.n, c.b), order=order, reorth="cgs2")
    runs = {}
    for dense in (False, True):
        out = solve(dA, c.b, prm, v0(ctx, c.csr.n, c.b), order=order, reorth="cgs2", rr_dense=dense)
        r = out.result
        print(f"  rr_dense={dense}: {r['iterations']} restarts, {r['rr_steps']} RR steps ({r['rr_structured']} structured), "
              f"host {r['host_seconds'] * 1e3 / max(r['rr_steps'], 1):.2f} ms per step, solve {r['seconds']:.3f} s", flush=True)
        assert r["converged"] == 1
        assert out.resid.max() <= c.tol
        runs[dense] = out
    s_, d_ = runs[False].result, runs[True].result
    assert s_["rr_structured"] == s_["rr_steps"] > 0
    assert d_["rr_structured"] == 0
    # both FP64 eigensolvers are accurate to eps ||T|| in absolute terms
    # (||T|| <= ||A|| = 8 here), which for the smallest Ritz value (1.2e-3)
    # is about 1e-12 relative; the two runs agree to that through the solve
    diff = np.abs(runs[False].evals - runs[True].evals)
    print("  max abs eigenvalue difference", diff.max(), "relative", (diff / np.abs(runs[True].evals)).max())
    assert diff.max() <= 1e-12 * 8
    if order == "ascending":
        assert np.max(np.abs(runs[False].evals - c.exact) / c.exact) <= 1e-10


def test_matrix_market_file_solve(ctx, tmp_path):
    # the 128^2 Laplacian written as a symmetric Matrix Market file (lower
    # triangle, 1-based) and read back by the library's loader: the solve
    # gives the closed-form eigenvalues, as from the in-memory CSR
    import scipy.io

    from trl_amd.formats import read_matrix_market

    c = P.config_c1()
    f = tmp_path / "laplace2d_128.mtx"
    scipy.io.mmwrite(str(f), c.csr.to_scipy(), symmetry="symmetric", precision=17)
    A = read_matrix_market(f)
    assert A.digest() == c.csr.digest()
    dA = DeviceCSR(ctx, A, precisions=("f64",))
    out = solve(dA, c.b, EigensolverParams(c.nev, c.ncv, tolerance=c.tol), v0(ctx, A.n, c.b), order="ascending", reorth="dgks")
    assert out.result["converged"] == 1
    assert np.abs(out.evals - c.exact).max() <= 1e-10 * c.exact.max()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_tuned_operator_solve(ctx, dtype):
    # the autotuned operator (CSR variants or ELL) gives the same solve
    c = P.config_c1()
    print(f"tuned operator {dtype}", flush=True)
    dA = DeviceCSR(ctx, c.csr)
    dt = np.float64 if dtype == "f64" else np.float32
    tol = c.tol if dtype == "f64" else 1e-5
    o = solve(dA, c.b, EigensolverParams(c.nev, c.ncv, tolerance=tol), v0(ctx, c.csr.n, c.b, dt), dtype=dtype, order="ascending", spmm_tune=True)
    assert o.result["converged"] == 1
    assert np.all(np.abs(o.evals - c.exact) <= (1e-10 if dtype == "f64" else 1e-4) * c.exact)
    om = solve_mixed(dA, c.b, EigensolverParams(c.nev, c.ncv, tolerance=c.tol), v0(ctx, c.csr.n, c.b), tol_fp32=1e-3, spmm_tune=True)
    assert om.result["converged"] == 1
    assert np.all(np.abs(om.evals - c.exact) <= 1e-10 * c.exact)


@pytest.mark.parametrize("reorth", ["cgs2", "dgks"])
def test_generalized_matches_scipy_golden(ctx, reorth):
    import json
    from pathlib import Path

    g = json.loads((Path(__file__).parent / "golden" / "stiffness_mass_eigsh.json").read_text())
    c = P.config_c2()
    assert c.csr.digest() == g["csr_sha256"]
    dA = DeviceCSR(ctx, c.csr, mass=c.mass, precisions=("f64",))
    out = solve(dA, c.b, EigensolverParams(c.nev, c.ncv), v0(ctx, c.csr.n, c.b), order="ascending", reorth=reorth)
    assert out.result["converged"] == 1
    ref = np.array(g["eigenvalues"])
    print("rel vs scipy", (np.abs(out.evals - ref) / ref).max())
    assert np.all(np.abs(out.evals - ref) <= 1e-10 * ref)
    # The solve runs on M^{-1} K in the M-inner product, and its stopping
    # estimate ||beta y_last|| / |theta| < tol (evp_base.hh:125-151) bounds
    # ||M^{-1} K x - theta x||_M / ||M^{-1} K x||_M = ||K x - theta M x||_{M^-1}
    # / ||K x||_{M^-1}: that residual meets the tolerance.  The 2-norm
    # residual the library reports is bounded only up to sqrt(cond(M)).
    X = np.concatenate(list(out.X.download()), axis=1)[:, : c.nev]
    KX = c.csr.to_scipy() @ X
    R = KX - (c.mass[:, None] * X) * out.evals
    w = 1.0 / np.sqrt(c.mass)[:, None]
    rm = np.linalg.norm(R * w, axis=0) / np.linalg.norm(KX * w, axis=0)
    print("M^-1-norm residuals", np.array2string(rm, precision=2), "2-norm residuals", np.array2string(out.resid, precision=2),
          "sqrt(cond(M))", float(np.sqrt(c.mass.max() / c.mass.min())))
    assert rm.max() <= c.tol
    assert out.resid.max() <= c.tol * np.sqrt(c.mass.max() / c.mass.min())


def test_fp32_and_mixed_laplace2d(ctx):
    c = P.config_c1()
    dA = DeviceCSR(ctx, c.csr)
    o32 = solve(dA, c.b, EigensolverParams(c.nev, c.ncv, tolerance=1e-5), v0(ctx, c.csr.n, c.b, np.float32), dtype="f32", order="ascending")
    assert o32.result["converged"] == 1
    assert np.all(np.abs(o32.evals - c.exact) <= 1e-4 * c.exact)
    om = solve_mixed(dA, c.b, EigensolverParams(c.nev, c.ncv, tolerance=1e-8), v0(ctx, c.csr.n, c.b), tol_fp32=1e-5)
    print("mixed", om.result, om.result32, (np.abs(om.evals - c.exact) / c.exact).max())
    assert om.result["converged"] == 1
    assert np.all(np.abs(om.evals - c.exact) <= 1e-10 * c.exact)
    assert om.resid.max() <= 1e-8


def test_errors(ctx):
    A = P.diagonal(64)
    dA = DeviceCSR(ctx, A, precisions=("f64",))
    with pytest.raises(L.TrlHipError) as e:
        solve(dA, 4, EigensolverParams(8, 64), v0(ctx, 64, 4))  # ncv + b > n (lanczos.hh:45-48)
    assert e.value.code == -1
    with pytest.raises(L.TrlHipError) as e:
        solve(dA, 4, EigensolverParams(8, 8), v0(ctx, 64, 4))  # nev >= ncv
    assert e.value.code == -1
    V = P.initial_block(64, 4)
    V[:, 1] = 0.0  # rank-deficient start block: Cholesky breakdown (evp_base.hh:55)
    with pytest.raises(L.TrlHipError) as e:
        solve(dA, 4, EigensolverParams(8, 32), ctx.to_device(V))
    assert e.value.code == -2
    assert ctx.check() == 0


def test_max_restarts_not_converged(ctx):
    c = P.config_c1()
    dA = DeviceCSR(ctx, c.csr, precisions=("f64",))
    out = solve(dA, c.b, EigensolverParams(c.nev, 32, max_restarts=2), v0(ctx, c.csr.n, c.b), order="ascending")
    assert out.result["converged"] == 0 and out.result["iterations"] == 2


@pytest.mark.parametrize("reorth", ["cgs2", "dgks"])
@pytest.mark.parametrize("phase1_fp64", [False, True])
def test_mixed_refine_laplace3d(ctx, phase1_fp64, reorth):
    # 3D Laplacian 24^3 + diag U[0,1] (a small stand-in for 160^3), b = 8, 32 smallest:
    # FP32 block Lanczos to 1e-3, then FP64 Chebyshev/Rayleigh-Ritz sweeps; the
    # result must match the plain FP64 solve and the CPU oracle
    c = P.config_c3(24)
    dA = DeviceCSR(ctx, c.csr)
    V0 = v0(ctx, c.csr.n, c.b)
    prm = EigensolverParams(c.nev, c.ncv, max_restarts=500)
    o64 = solve(dA, c.b, prm, V0, reorth=reorth)
    om = solve_mixed(dA, c.b, prm, V0, tol_fp32=1e-3, cheb_degree=60, phase1_fp64=phase1_fp64, reorth=reorth)
    print("mixed", om.result, om.result32)
    assert om.result["converged"] == 1
    assert om.resid.max() <= 1e-8
    assert np.all(np.abs(om.evals - o64.evals) <= 1e-10 * np.abs(o64.evals))
    ref = oracle.lanczos_solve(c.csr, c.b, c.nev, c.ncv, P.initial_block(c.csr.n, c.b), order=1, ritz_vectors=True)
    assert np.all(np.abs(om.evals - ref["evals"]) <= 1e-10 * np.abs(ref["evals"]))
    Xh = np.concatenate(list(om.X.download()), axis=1)[:, : c.nev]
    print("max principal angle vs oracle", max_angle(Xh, ref["X"]))
    assert max_angle(Xh, ref["X"]) < 1e-6


@pytest.mark.parametrize("degree_opt", [0, 1, 2])
@pytest.mark.parametrize("residual_filter", [False, True])
def test_mixed_fp32_refinement_sweeps(ctx, degree_opt, residual_filter):
    # FP32 Lanczos to a loose 5e-2, FP32 Chebyshev/Rayleigh-Ritz sweeps down to
    # the FP32 floor, then FP64 sweeps (every block at the full degree, or each
    # block at the degree its convergence rate asks for, in every sweep or in
    # the FP64 ones): the same eigenpairs as the plain FP64 solve
    c = P.config_c3(24)
    dA = DeviceCSR(ctx, c.csr)
    V0 = v0(ctx, c.csr.n, c.b)
    prm = EigensolverParams(c.nev, c.ncv, max_restarts=500)
    o64 = solve(dA, c.b, prm, V0, reorth="dgks")
    om = solve_mixed(dA, c.b, prm, V0, tol_fp32=5e-2, cheb_degree=60, refine_blocks=9, max_refine=14, fp32_sweeps=8, tol_fp32_refine=3e-5,
                     reorth="dgks", spmm_tune=True, degree_opt=degree_opt, residual_filter=residual_filter)
    print("mixed with fp32 sweeps", degree_opt, residual_filter, om.result, om.result32)
    assert om.result["converged"] == 1
    assert om.resid.max() <= 1e-8
    assert np.all(np.abs(om.evals - o64.evals) <= 1e-10 * np.abs(o64.evals))


def test_mixed_filter_all_blocks_per_launch(ctx):
    # a 3D Laplacian large enough that stepping all 9 refinement blocks in one
    # launch tunes faster than 9 launches (the solver then filters that way):
    # the same eigenpairs as the plain FP64 solve
    c = P.config_c3(64)
    dA = DeviceCSR(ctx, c.csr)
    op = dA.operator(c.b, "f32", tune=True)
    op.tune_panels(9)
    print("panels", op.name_axpby, op.us_axpby, op._name(op.desc_panels), op.us_panels, op.panels_pay(9))
    if not op.panels_pay(9):
        pytest.skip("one launch per block tuned faster on this device")
    V0 = v0(ctx, c.csr.n, c.b)
    prm = EigensolverParams(c.nev, c.ncv, max_restarts=500)
    o64 = solve(dA, c.b, prm, V0, reorth="dgks", spmm_tune=True)
    om = solve_mixed(dA, c.b, prm, V0, tol_fp32=5e-2, cheb_degree=80, refine_blocks=9, max_refine=14, fp32_sweeps=8, tol_fp32_refine=3e-5,
                     reorth="dgks", spmm_tune=True)
    print("mixed", om.result, om.result32)
    assert om.result["converged"] == 1
    assert om.resid.max() <= 1e-8
    assert np.all(np.abs(om.evals - o64.evals) <= 1e-10 * np.abs(o64.evals))


@pytest.mark.parametrize("residual_filter", [False, True])
def test_mixed_bench_schedule(ctx, residual_filter):
    # the benchmark's schedule on a 48^3 stand-in: one loose FP32 Lanczos
    # cycle over 192 columns, FP32 sweeps on 5 blocks, FP64 sweeps, with
    # per-block degrees (the FP64 sweeps' filter in FP64, or in FP32 driven
    # by the FP64 residuals)
    c = P.config_c3(48)
    dA = DeviceCSR(ctx, c.csr)
    V0 = v0(ctx, c.csr.n, c.b)
    prm = EigensolverParams(c.nev, c.ncv, max_restarts=500)
    o64 = solve(dA, c.b, prm, V0, reorth="dgks", spmm_tune=True)
    om = solve_mixed(dA, c.b, prm, V0, tol_fp32=0.4, cheb_degree=160, refine_blocks=5, max_refine=14, fp32_sweeps=12, tol_fp32_refine=1e-3,
                     reorth="dgks", spmm_tune=True, degree_opt=1, residual_filter=residual_filter, ncv_fp32=192)
    print("bench schedule", residual_filter, om.result, om.result32)
    assert om.result["converged"] == 1
    assert om.resid.max() <= 1e-8
    assert np.all(np.abs(om.evals - o64.evals) <= 1e-10 * np.abs(o64.evals))


def test_mixed_generalized(ctx):
    # the structural pair K x = lambda M x in mixed precision, against scipy
    import json
    from pathlib import Path

    g = json.loads((Path(__file__).parent / "golden" / "stiffness_mass_eigsh.json").read_text())
    c = P.config_c2()
    dA = DeviceCSR(ctx, c.csr, mass=c.mass)
    om = solve_mixed(dA, c.b, EigensolverParams(c.nev, c.ncv), v0(ctx, c.csr.n, c.b), tol_fp32=1e-4, cheb_degree=40)
    ref = np.array(g["eigenvalues"])
    print("mixed generalized", om.result, om.result32, (np.abs(om.evals - ref) / ref).max())
    assert np.all(np.abs(om.evals - ref) <= 1e-10 * ref)
    assert om.resid.max() <= 1e-8
    # FP32 sweeps, then the FP64 sweeps' filter in FP32 on M^{-1} K driven by
    # the FP64 residuals
    om = solve_mixed(dA, c.b, EigensolverParams(c.nev, c.ncv), v0(ctx, c.csr.n, c.b), tol_fp32=1e-2, cheb_degree=40, max_refine=14, fp32_sweeps=6,
                     tol_fp32_refine=3e-5, residual_filter=True)
    print("mixed generalized, residual filter", om.result, om.result32, (np.abs(om.evals - ref) / ref).max())
    assert om.result["converged"] == 1
    assert np.all(np.abs(om.evals - ref) <= 1e-10 * ref)
    assert om.resid.max() <= 1e-8


@pytest.mark.parametrize("reorth", ["mgs", "dgks"])
def test_mixed_thick_restart_warm_start(ctx, reorth):
    # refine = 2: the FP32 solve's thick-restart state (Ritz blocks, the next
    # block, diag(theta) + arrow) carried to FP64 and continued by the FP64
    # driver's cycles (BlockLanczos::solve_restarted).  The eigenvalues match
    # the plain FP64 solve; the true residuals keep the FP32 basis' rounding
    # (about 1e-5 here), which the FP64 cycles do not see.
    c = P.config_c3(24)
    dA = DeviceCSR(ctx, c.csr)
    V0 = v0(ctx, c.csr.n, c.b)
    prm = EigensolverParams(c.nev, c.ncv, max_restarts=500)
    o64 = solve(dA, c.b, prm, V0, reorth=reorth)
    for tol32 in (1e-3, 3e-5):
        om = solve_mixed(dA, c.b, prm, V0, refine=2, tol_fp32=tol32, reorth=reorth)
        rel = np.abs(om.evals - o64.evals) / np.abs(o64.evals)
        print(f"warm start {reorth} tol32={tol32}: fp32 {om.result32['iterations']} cycles, fp64 {om.result['iterations']} cycles, "
              f"max rel eig {rel.max():.2e}, true residual {om.resid.max():.2e}")
        assert om.result["converged"] == 1 and om.result32["converged"] == 1
        assert rel.max() <= 1e-9
        assert om.resid.max() <= 1e-4
