"""The allocation-free Gram entry point (trl_hip_gram_noalloc) against
trl_hip_gram, through the C-ABI: two calls in a row, the second on a smaller
input, FP32 and FP64, with and without a mass diagonal; and its size limit.
Bar: within 1e-14 relative of trl_hip_gram in FP64 and 1e-6 in FP32 (the same
kernels; only the workgroup count, and with it the summation order, may
differ where the fixed scratch caps it), deterministic across repeated calls."""

import numpy as np
import pytest

from trl_amd import _lib as L

pytestmark = pytest.mark.gpu

DT = {"f64": np.float64, "f32": np.float32}


def gram(ctx, fn, t, n, b, k, V, W, m):
    C = ctx.to_device(np.full(k * b * b, np.nan))
    L.check(fn(ctx.h, L.dtype_code(DT[t]), n, b, k, V.ptr, W.ptr, m.ptr if m is not None else None, C.ptr), "gram")
    return C.download()


@pytest.mark.parametrize("t", ["f64", "f32"])
@pytest.mark.parametrize("mass", [False, True])
def test_gram_noalloc_matches_gram(ctx, t, mass):
    H = L.hip()
    rng = np.random.default_rng(7)
    dt = DT[t]
    # (n, b, k): a large call, then a smaller one; the first at the largest
    # coefficient count the entry point takes (k b^2 = 4096)
    for n, b, k in [(300_001, 8, 64), (70_003, 8, 16), (1_000_000, 4, 1), (5_000, 6, 3)]:
        print(f"gram_noalloc {t} mass={mass} n={n} b={b} k={k}", flush=True)
        V = ctx.to_device(rng.standard_normal((k, n, b)).astype(dt))
        W = ctx.to_device(rng.standard_normal((n, b)).astype(dt))
        m = ctx.to_device(rng.uniform(0.5, 2.0, n).astype(dt)) if mass else None
        ref = gram(ctx, H.trl_hip_gram, t, n, b, k, V, W, m)
        got = gram(ctx, H.trl_hip_gram_noalloc, t, n, b, k, V, W, m)
        again = gram(ctx, H.trl_hip_gram_noalloc, t, n, b, k, V, W, m)
        assert np.array_equal(got, again)
        r = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"  rel diff vs trl_hip_gram {r:.2e}, bitwise equal {np.array_equal(got, ref)}", flush=True)
        assert r <= (1e-14 if t == "f64" else 1e-6)


def test_gram_noalloc_limits(ctx):
    H = L.hip()
    n, b = 1000, 8
    V = ctx.to_device(np.ones((65, n, b)))
    W = ctx.to_device(np.ones((n, b)))
    C = ctx.empty((65 * b * b,))
    # k b^2 = 4160 > TRL_HIP_NOALLOC_MAX_COEFS: refused
    assert H.trl_hip_gram_noalloc(ctx.h, L.F64, n, b, 65, V.ptr, W.ptr, None, C.ptr) == -1
    assert H.trl_hip_gram_noalloc(ctx.h, L.F64, n, 5, 1, V.ptr, W.ptr, None, C.ptr) == -1
    assert H.trl_hip_gram_noalloc(None, L.F64, n, b, 1, V.ptr, W.ptr, None, C.ptr) == -1
    # n = 0: a zero Gram matrix
    L.check(H.trl_hip_gram_noalloc(ctx.h, L.F64, 0, b, 2, V.ptr, W.ptr, None, C.ptr), "n=0")
    assert np.all(C.download()[: 2 * b * b] == 0.0)
