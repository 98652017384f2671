"""The context's allocation cache (TRL_HIP_OPT_ALLOC_CACHE).

trl_hip_free keeps a block of trl_hip_malloc for the next allocation of the
same size: the block comes back zero-filled (the allocation's contract,
multivector.hh:44-48), other sizes get blocks of their own, and with the
option off the cache is emptied and blocks are released on free.
"""

import ctypes as C

import numpy as np
import pytest

from trl_amd import _lib as L

pytestmark = pytest.mark.gpu


def malloc(ctx, nbytes):
    p = C.c_void_p()
    L.check(L.hip().trl_hip_malloc(ctx.h, nbytes, C.byref(p)), "malloc")
    return p.value


def test_alloc_cache_reuses_zeroed_blocks():
    H = L.hip()
    ctx = L.Context(0)
    try:
        v = C.c_int(-1)
        L.check(H.trl_hip_ctx_get_option(ctx.h, L.OPT_ALLOC_CACHE, C.byref(v)), "get option")
        assert v.value == 1
        nbytes = 3 << 20
        host = np.arange(nbytes // 8, dtype=np.float64) + 1.0
        p = malloc(ctx, nbytes)
        L.check(H.trl_hip_memcpy_h2d(ctx.h, C.c_void_p(p), host.ctypes.data, nbytes), "upload")
        L.check(H.trl_hip_free(ctx.h, C.c_void_p(p)), "free")
        q = malloc(ctx, nbytes)  # the cached block, zero-filled again
        assert q == p
        back = np.empty_like(host)
        L.check(H.trl_hip_memcpy_d2h(ctx.h, back.ctypes.data, C.c_void_p(q), nbytes), "download")
        assert not back.any()
        r = malloc(ctx, nbytes + 4096)  # another size: a block of its own
        assert r not in (p, q)
        for a in (q, r):
            L.check(H.trl_hip_free(ctx.h, C.c_void_p(a)), "free")
        # the option off empties the cache; blocks are then released on free
        L.check(H.trl_hip_ctx_set_option(ctx.h, L.OPT_ALLOC_CACHE, 0), "option off")
        s = malloc(ctx, nbytes)
        L.check(H.trl_hip_memcpy_d2h(ctx.h, back.ctypes.data, C.c_void_p(s), nbytes), "download")
        assert not back.any()
        L.check(H.trl_hip_free(ctx.h, C.c_void_p(s)), "free")
        L.check(H.trl_hip_ctx_get_option(ctx.h, L.OPT_ALLOC_CACHE, C.byref(v)), "get option")
        assert v.value == 0
    finally:
        ctx.close()
