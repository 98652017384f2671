"""Batched lock-step solves (trl_hip_batch_solve_csr) against single-problem
FP64 solves of the same problems from the same start blocks, against the CPU
oracle, and against themselves (graph replay vs direct launches, kernel
variants).  Bar: eigenvalues within 1e-10 relative, true residuals at the
solver tolerance."""

import numpy as np
import pytest

import oracle
from trl_amd import problems as P
from trl_amd.batch import solve_batch
from trl_amd.solver import DeviceCSR, EigensolverParams, solve

pytestmark = pytest.mark.gpu

B, NEV, NCV = 4, 10, 48


def small_batch(sides=(20, 33, 24, 41, 28, 36)):
    mats = []
    for p, s in enumerate(sides):
        pot = 2.0 * P.uniform(7, 1, s * s, offset=1000 * p)
        mats.append(P.laplacian_2d(s, s, pot))
    V0 = [P.initial_block(m.n, B, seed=42 + p) for p, m in enumerate(mats)]
    return mats, V0


def single(ctx, A, V0, nev=NEV, ncv=NCV):
    dA = DeviceCSR(ctx, A, precisions=("f64",))
    return solve(dA, B, EigensolverParams(nev, ncv), ctx.to_device(V0), order="ascending", reorth="cgs2")


@pytest.mark.parametrize("tune", [False, True])
def test_batch_matches_single_problem_solves(ctx, tune):
    print(f"batch small tune={tune}", flush=True)
    mats, V0 = small_batch()
    out = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=tune, ritz_vectors=True)
    print("stats", out.stats)
    for p, (A, v0) in enumerate(zip(mats, V0)):
        r = out.results[p]
        assert r["converged"] == 1, (p, r)
        ref = single(ctx, A, v0)
        rel = np.abs(out.evals[p] - ref.evals) / np.abs(ref.evals)
        print(f"problem {p} n={A.n}: cycles {r['iterations']} (single {ref.result['iterations']}), max rel {rel.max():.2e}, "
              f"resid {out.resid[p].max():.2e}")
        assert rel.max() <= 1e-10
        assert out.resid[p].max() <= 1e-8  # the solve's tolerance
        # Ritz vectors: orthonormal and A x = lambda x
        Xh = np.concatenate(list(out.X[p]), axis=1)[:, :NEV]
        assert np.abs(Xh.T @ Xh - np.eye(NEV)).max() < 1e-10
        AX = A.to_scipy() @ Xh
        assert np.abs(AX - Xh * out.evals[p]).max() <= 1e-7 * np.abs(AX).max()


def test_batch_matches_oracle(ctx):
    mats, V0 = small_batch((26, 31))
    out = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=False)
    for p, A in enumerate(mats):
        ref = oracle.lanczos_solve(A, B, NEV, NCV, V0[p], order=1)
        assert ref["converged"]
        rel = np.abs(out.evals[p] - ref["evals"]) / np.abs(ref["evals"])
        print(f"problem {p}: max rel vs oracle {rel.max():.2e}")
        assert rel.max() <= 1e-10


def test_batch_graph_replay_is_bitwise_eager(ctx):
    # the captured graphs replay the same deterministic kernels as direct launches
    mats, V0 = small_batch((22, 30, 25))
    g = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=False, use_graph=True)
    e = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=False, use_graph=False)
    assert np.array_equal(g.evals, e.evals)
    assert [r["iterations"] for r in g.results] == [r["iterations"] for r in e.results]
    assert g.stats["cycles"] >= 2  # graphs were replayed across restarts


def test_batch_adaptive_reorthogonalisation(ctx):
    # the second Gram-Schmidt pass only where the DGKS test asks: the same
    # eigenpairs as single-problem solves with the same test, and as the
    # two-pass batch; graph replay still bitwise equal to direct launches
    mats, V0 = small_batch((23, 37, 29, 44))
    out = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=True, reorth="dgks", ritz_vectors=True)
    two = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=True, reorth="cgs2")
    for p, (A, v0) in enumerate(zip(mats, V0)):
        assert out.results[p]["converged"] == 1
        dA = DeviceCSR(ctx, A, precisions=("f64",))
        ref = solve(dA, B, EigensolverParams(NEV, NCV), ctx.to_device(v0), order="ascending", reorth="dgks")
        rel = np.abs(out.evals[p] - ref.evals) / np.abs(ref.evals)
        rel2 = np.abs(out.evals[p] - two.evals[p]) / np.abs(two.evals[p])
        print(f"problem {p} n={A.n}: cycles {out.results[p]['iterations']} (single {ref.result['iterations']}, two-pass batch "
              f"{two.results[p]['iterations']}), max rel {rel.max():.2e} / {rel2.max():.2e}")
        assert rel.max() <= 1e-10 and rel2.max() <= 1e-10
        assert out.resid[p].max() <= 1e-8  # the solve's tolerance
        Xh = np.concatenate(list(out.X[p]), axis=1)[:, :NEV]
        assert np.abs(Xh.T @ Xh - np.eye(NEV)).max() < 1e-10
    e = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=False, reorth="dgks", use_graph=False)
    g = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=False, reorth="dgks", use_graph=True)
    assert np.array_equal(g.evals, e.evals)


def test_batch_cholqr2_and_errors(ctx):
    from trl_amd import _lib as L

    mats, V0 = small_batch((21, 27))
    o1 = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=False, cholqr_passes=2)
    o0 = solve_batch(ctx, mats, B, NEV, NCV, V0, tune=False)
    assert np.all(np.abs(o1.evals - o0.evals) <= 1e-10 * np.abs(o0.evals))
    tiny = P.laplacian_2d(6, 6)  # n = 36 < ncv + b
    with pytest.raises(L.TrlHipError) as e:
        solve_batch(ctx, [mats[0], tiny], B, NEV, NCV, [V0[0], P.initial_block(36, B)])
    assert e.value.code == -1


def test_batch_c4_64_uneven_problems(ctx):
    # 64 2D Laplacians, sides U{48..256}, potential U[0, 2], b = 4, 10 smallest, ncv = 128
    cfgs = P.config_c4()
    mats = [c.csr for c in cfgs]
    ncv = cfgs[0].ncv
    V0 = [P.initial_block(m.n, B, seed=42 + p) for p, m in enumerate(mats)]
    print(f"c4: {len(mats)} problems, n {min(m.n for m in mats)}..{max(m.n for m in mats)}, ncv {ncv}", flush=True)
    out = solve_batch(ctx, mats, B, NEV, ncv, V0, tune=True)
    st = out.stats
    t = np.array([r["seconds"] for r in out.results])
    print(f"batch {st['seconds']:.3f} s, {len(mats) / st['seconds']:.1f} problems/s, p50 {np.percentile(t, 50) * 1e3:.1f} ms, "
          f"p99 {np.percentile(t, 99) * 1e3:.1f} ms, cycles {st['cycles']}, host {st['host_seconds']:.3f} s, tiles {st['ntiles']}, "
          f"format {st['format']}, rows/thread {st['rows_per_thread']}")
    assert all(r["converged"] == 1 for r in out.results)
    assert out.resid.max() <= 1e-8  # the solve's tolerance
    # spot checks against single-problem solves: the smallest, the largest and two others
    order = np.argsort([m.n for m in mats])
    for p in (order[0], order[-1], order[20], order[45]):
        ref = single(ctx, mats[p], V0[p], ncv=ncv)
        rel = np.abs(out.evals[p] - ref.evals) / np.abs(ref.evals)
        print(f"problem {p} n={mats[p].n}: max rel vs single {rel.max():.2e}")
        assert rel.max() <= 1e-10


def test_batch_barrier_error_path(ctx):
    # the barrier error flag raised from the host, as a timed-out device
    # barrier raises it: trl_hip_batch_extend returns TRL_HIP_EBARRIER (every
    # step kernel returns at once, so every launch ends); after
    # trl_hip_batch_reset_flags the same extend runs and gives the
    # coefficients of an undisturbed run bit for bit
    import ctypes as C

    from trl_amd import _lib as L

    H = L.hip()
    mats, V0 = small_batch((18, 23))
    nprob, m = len(mats), NCV // B
    keep = [np.ascontiguousarray(a) for A in mats for a in (A.rowptr.astype(np.int32), A.col.astype(np.int32), A.val.astype(np.float64))]
    n = (C.c_size_t * nprob)(*[A.n for A in mats])
    rp = (C.c_void_p * nprob)(*[keep[3 * p].ctypes.data for p in range(nprob)])
    ci = (C.c_void_p * nprob)(*[keep[3 * p + 1].ctypes.data for p in range(nprob)])
    va = (C.c_void_p * nprob)(*[keep[3 * p + 2].ctypes.data for p in range(nprob)])
    bt = C.c_void_p()
    L.check(H.trl_hip_batch_create(ctx.h, nprob, n, rp, ci, va, B, NCV, C.byref(bt)), "batch_create")
    try:
        def extend():
            for p in range(nprob):
                v = np.ascontiguousarray(V0[p])
                L.check(H.trl_hip_batch_set_block(bt, p, 0, v.ctypes.data), "set_block")
            alpha = np.zeros(nprob * m * B * B)
            beta = np.zeros(nprob * m * B * B)
            pflag = np.zeros(nprob, dtype=np.int32)
            rc = H.trl_hip_batch_extend(bt, 0, m, 0, None, alpha.ctypes.data, beta.ctypes.data, pflag.ctypes.data)
            return rc, alpha, beta, pflag

        rc, a0, b0, f0 = extend()
        assert rc == 0 and not f0.any()
        L.check(H.trl_hip_batch_raise_error(bt), "raise_error")
        rc, _, _, _ = extend()
        print("extend with the barrier error flag raised:", rc, L.hip().trl_hip_strerror(rc))
        assert rc == -6
        rc, _, _, _ = extend()  # the flag is sticky until reset
        assert rc == -6
        L.check(H.trl_hip_batch_reset_flags(bt), "reset_flags")
        rc, a1, b1, f1 = extend()
        assert rc == 0 and not f1.any()
        assert np.array_equal(a0, a1) and np.array_equal(b0, b1)
    finally:
        H.trl_hip_batch_destroy(bt)
