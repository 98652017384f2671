"""ctypes binding of the C-ABI in include/trl_hip.h and include/trl_hip_solver.h.

Loading fails loudly: there is no CPU fallback for any operation of this
package.  The libraries are the in-tree builds under trl_amd/lib/.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

F32 = 0
F64 = 1
# trl_hip_ctx_set_option options
OPT_OFFSETS64 = 1
OPT_GRAPHS = 2
OPT_ALLOC_CACHE = 3

c_size = C.c_size_t
c_int = C.c_int
c_void_p = C.c_void_p
c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)



class SpMat(C.Structure):
    """trl_hip_spmat: a device-resident sparse operator and its kernel variant.

    format 0 = CSR (rowptr, col, val), 1 = slot-major ELL (width, col, val);
    plane / plane_block: blocked row order (0 = index order); share, line,
    tile_lines: options of the several-panel ELL step."""

    _fields_ = [("format", c_int), ("rowptr", c_void_p), ("width", c_int), ("col", c_void_p), ("val", c_void_p), ("dscale", c_void_p),
                ("variant", c_int), ("plane", c_int), ("plane_block", c_int), ("share", c_int), ("line", c_int), ("tile_lines", c_int)]


c_spmat_p = C.POINTER(SpMat)

# (name, restype, argtypes) of every symbol declared in include/trl_hip.h
HIP_API = [
    ("trl_hip_ctx_create", c_int, [c_int, C.POINTER(c_void_p)]),
    ("trl_hip_ctx_destroy", c_int, [c_void_p]),
    ("trl_hip_ctx_sync", c_int, [c_void_p]),
    ("trl_hip_ctx_check", c_int, [c_void_p, c_int_p]),
    ("trl_hip_ctx_stream", c_void_p, [c_void_p]),
    ("trl_hip_ctx_num_cus", c_int, [c_void_p]),
    ("trl_hip_ctx_set_option", c_int, [c_void_p, c_int, c_int]),
    ("trl_hip_ctx_get_option", c_int, [c_void_p, c_int, c_int_p]),
    ("trl_hip_strerror", C.c_char_p, [c_int]),
    ("trl_hip_malloc", c_int, [c_void_p, c_size, C.POINTER(c_void_p)]),
    ("trl_hip_free", c_int, [c_void_p, c_void_p]),
    ("trl_hip_memcpy_h2d", c_int, [c_void_p, c_void_p, c_void_p, c_size]),
    ("trl_hip_memcpy_d2h", c_int, [c_void_p, c_void_p, c_void_p, c_size]),
    ("trl_hip_memcpy_d2d", c_int, [c_void_p, c_void_p, c_void_p, c_size]),
    ("trl_hip_fill_zero", c_int, [c_void_p, c_void_p, c_size]),
    ("trl_hip_convert", c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_size]),
    ("trl_hip_panel_sub", c_int, [c_void_p, c_int, c_size, c_int, c_void_p, c_void_p]),
    ("trl_hip_panel_mult", c_int, [c_void_p, c_int, c_size, c_int, c_void_p, c_void_p, c_void_p, c_int]),
    ("trl_hip_gram", c_int, [c_void_p, c_int, c_size, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_gram_noalloc", c_int, [c_void_p, c_int, c_size, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_update", c_int, [c_void_p, c_int, c_size, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_cholqr", c_int, [c_void_p, c_int, c_size, c_int, c_void_p, c_void_p, c_void_p, c_int]),
    ("trl_hip_graph_begin", c_int, [c_void_p]),
    ("trl_hip_graph_end", c_int, [c_void_p, C.POINTER(c_void_p)]),
    ("trl_hip_graph_launch", c_int, [c_void_p, c_void_p]),
    ("trl_hip_graph_destroy", c_int, [c_void_p]),
    ("trl_hip_cholqr_bprod", c_int, [c_void_p, c_int, c_size, c_int, c_void_p, c_void_p, c_void_p, c_int]),
    ("trl_hip_pcg_workspace", c_size, [c_size, c_int]),
    ("trl_hip_pcg", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_void_p, c_void_p, c_void_p, C.c_double, c_int, c_void_p, c_int_p, c_double_p]),
    ("trl_hip_rotate", c_int, [c_void_p, c_int, c_size, c_int, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    ("trl_hip_spmm", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_void_p, c_void_p]),
    ("trl_hip_spmm_axpby", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_void_p, c_void_p, c_void_p, C.c_double, C.c_double, C.c_double]),
    ("trl_hip_spmm_panels", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_int, c_void_p, c_void_p]),
    ("trl_hip_spmm_axpby_panels", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_int, c_void_p, c_void_p, c_void_p, C.c_double, C.c_double,
                                          C.c_double]),
    ("trl_hip_spmm_axpby_forced", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_int, c_void_p, c_void_p, c_void_p, C.c_double, C.c_double,
                                          C.c_double, c_void_p, c_void_p]),
    ("trl_hip_residual_split", c_int, [c_void_p, c_size, c_int, c_int, c_void_p, c_void_p, c_void_p, C.c_double, c_void_p, c_void_p]),
    ("trl_hip_panel_combine", c_int, [c_void_p, c_size, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_csr_max_row", c_int, [c_void_p, c_size, c_void_p, c_int_p]),
    ("trl_hip_csr_to_ell", c_int, [c_void_p, c_int, c_size, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    ("trl_hip_spmm_tune", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_spmat_p, c_spmat_p, c_double_p]),
    ("trl_hip_spmm_tune_axpby", c_int, [c_void_p, c_int, c_size, c_int, c_spmat_p, c_spmat_p, c_spmat_p, c_double_p]),
    ("trl_hip_spmm_tune_panels", c_int, [c_void_p, c_int, c_size, c_int, c_int, c_spmat_p, c_spmat_p, c_spmat_p, c_double_p]),
    ("trl_hip_spmm_csr", c_int, [c_void_p, c_int, c_size, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    ("trl_hip_spmm_csr_axpby", c_int, [c_void_p, c_int, c_size, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.c_double,
                                       C.c_double, C.c_double]),
    ("trl_hip_spmm_ell", c_int, [c_void_p, c_int, c_size, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_small_copy", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    ("trl_hip_small_zero", c_int, [c_void_p, c_int, c_void_p]),
    ("trl_hip_small_set_diagonal", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("trl_hip_small_mult", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_residual_sq", c_int, [c_void_p, c_int, c_size, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_reorth", c_int, [c_void_p, c_int, c_size, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, C.c_double]),
    ("trl_hip_prof_read_dtype", c_int, [c_void_p, c_int, c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("trl_hip_prof_reorth", c_int, [c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    ("trl_hip_prof_enable", c_int, [c_void_p, c_int]),
    ("trl_hip_region_begin", c_int, [c_void_p]),
    ("trl_hip_region_end", c_int, [c_void_p, C.POINTER(C.c_double)]),
    ("trl_hip_prof_read", c_int, [c_void_p, c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("trl_hip_batch_create", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, C.POINTER(c_void_p)]),
    ("trl_hip_batch_destroy", c_int, [c_void_p]),
    ("trl_hip_batch_set_variant", c_int, [c_void_p, c_int, c_int]),
    ("trl_hip_batch_tune", c_int, [c_void_p, c_double_p]),
    ("trl_hip_batch_info", c_int, [c_void_p, c_int_p, c_int_p, c_int_p, c_double_p]),
    ("trl_hip_batch_set_options", c_int, [c_void_p, c_int, c_int]),
    ("trl_hip_batch_set_reorth", c_int, [c_void_p, c_int]),
    ("trl_hip_batch_set_block", c_int, [c_void_p, c_int, c_int, c_void_p]),
    ("trl_hip_batch_get_block", c_int, [c_void_p, c_int, c_int, c_void_p]),
    ("trl_hip_batch_set_active", c_int, [c_void_p, c_void_p]),
    ("trl_hip_batch_extend", c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_batch_rotate", c_int, [c_void_p, c_void_p, c_int]),
    ("trl_hip_batch_save_blocks", c_int, [c_void_p, c_int, c_int]),
    ("trl_hip_batch_residuals", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("trl_hip_batch_reset_flags", c_int, [c_void_p]),
    ("trl_hip_batch_raise_error", c_int, [c_void_p]),
    ("trl_hip_sddmm_adjoint", c_int, [c_void_p, c_int, c_size, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
]

_hip = None


class TrlHipError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} (code {code})")


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build the HIP extension first (python -m trl_amd.build); there is no CPU fallback")
    return C.CDLL(str(path), mode=os.RTLD_NOW | C.RTLD_GLOBAL)


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        lib = _load("libtrl_hip.so")
        for name, res, args in HIP_API:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _hip = lib
    return _hip


def strerror(code: int) -> str:
    try:
        return hip().trl_hip_strerror(code).decode()
    except Exception:  # library missing
        return f"error {code}"


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise TrlHipError(rc, what)


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return F32
    if dt == np.float64:
        return F64
    raise ValueError(f"unsupported dtype {dt}")


class DeviceArray:
    """A device allocation with a numpy-like shape/dtype (row-major)."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = c_void_p()
        check(hip().trl_hip_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)), "trl_hip_malloc")
        self.ptr = p.value

    def offset(self, elements: int) -> int:
        return self.ptr + int(elements) * self.dtype.itemsize

    def upload(self, host: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(host, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(hip().trl_hip_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "h2d")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(hip().trl_hip_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().trl_hip_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class Context:
    """One device + one in-order stream (trl_hip_ctx)."""

    def __init__(self, device: int = 0):
        h = c_void_p()
        check(hip().trl_hip_ctx_create(device, C.byref(h)), "trl_hip_ctx_create")
        self.h = h.value

    def close(self) -> None:
        if self.h:
            hip().trl_hip_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self) -> None:
        check(hip().trl_hip_ctx_sync(self.h), "sync")

    def check(self) -> int:
        piv = c_int(0)
        rc = hip().trl_hip_ctx_check(self.h, C.byref(piv))
        if rc not in (0, -2):
            check(rc, "check")
        return piv.value

    PROF_FAMILIES = ["spmm", "gram", "update", "cholqr", "rotate", "panel", "small", "other"]

    def prof_enable(self, on: bool = True) -> None:
        check(hip().trl_hip_prof_enable(self.h, 1 if on else 0), "prof_enable")

    def prof_read(self) -> dict:
        out = {}
        for f, name in enumerate(self.PROF_FAMILIES):
            n, s, b = C.c_longlong(), C.c_double(), C.c_double()
            check(hip().trl_hip_prof_read(self.h, f, C.byref(n), C.byref(s), C.byref(b)), "prof_read")
            out[name] = {"calls": n.value, "seconds": s.value, "bytes": b.value}
        return out

    def prof_read_dtype(self, dtype: int) -> dict:
        """prof_read restricted to calls on panels of one dtype (F32 / F64)."""
        out = {}
        for f, name in enumerate(self.PROF_FAMILIES):
            n, s, b = C.c_longlong(), C.c_double(), C.c_double()
            check(hip().trl_hip_prof_read_dtype(self.h, f, dtype, C.byref(n), C.byref(s), C.byref(b)), "prof_read_dtype")
            out[name] = {"calls": n.value, "seconds": s.value, "bytes": b.value}
        return out

    def time_region(self, fn, reps: int) -> float:
        """Mean device seconds per call of fn over reps calls, from one HIP
        event pair on the context stream around all of them."""
        check(hip().trl_hip_region_begin(self.h), "region_begin")
        for _ in range(reps):
            fn()
        s = C.c_double()
        check(hip().trl_hip_region_end(self.h, C.byref(s)), "region_end")
        return s.value / reps

    def prof_reorth(self) -> dict:
        """trl_hip_reorth calls with the second-pass test since prof_enable, and
        how many took the second pass."""
        c, s = C.c_longlong(), C.c_longlong()
        check(hip().trl_hip_prof_reorth(self.h, C.byref(c), C.byref(s)), "prof_reorth")
        return {"calls": c.value, "second_passes": s.value}

    def set_option(self, option: int, value: int) -> None:
        check(hip().trl_hip_ctx_set_option(self.h, int(option), int(value)), "ctx_set_option")

    @property
    def num_cus(self) -> int:
        return hip().trl_hip_ctx_num_cus(self.h)

    def empty(self, shape, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def to_device(self, host: np.ndarray, dtype=None) -> DeviceArray:
        host = np.asarray(host)
        d = DeviceArray(self, host.shape, dtype or host.dtype)
        return d.upload(host)


# ---------------------------------------------------------------------------
# include/trl_hip_solver.h


class SolverParams(C.Structure):
    _fields_ = [("b", c_int), ("nev", C.c_uint), ("ncv", C.c_uint), ("max_restarts", C.c_uint), ("tol", C.c_double), ("order", c_int),
                ("reorth", c_int), ("cholqr_passes", c_int), ("lanes_per_row", c_int),
                ("spmm_tune", c_int), ("rr_dense", c_int)]


class SolverResult(C.Structure):
    _fields_ = [("converged", c_int), ("iterations", C.c_uint), ("n_op_apply", C.c_uint), ("seconds", C.c_double), ("host_seconds", C.c_double),
                ("rr_steps", C.c_uint), ("rr_structured", C.c_uint), ("graph_cycles", C.c_uint)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class MixedParams(C.Structure):
    _fields_ = [("tol_fp32", C.c_double), ("max_restarts_fp32", C.c_uint), ("seed", C.c_ulonglong), ("refine", c_int), ("cheb_degree", C.c_uint),
                ("max_refine", C.c_uint), ("phase1_fp64", c_int), ("refine_blocks", C.c_uint), ("fp32_sweeps", C.c_uint), ("tol_fp32_refine", C.c_double),
                ("degree_opt", c_int), ("residual_filter", c_int), ("ncv_fp32", C.c_uint)]


class InnerStats(C.Structure):
    """trl_hip_inner_stats: the shift-invert solve's conjugate-gradient counts."""

    _fields_ = [("solves", C.c_uint), ("iterations", C.c_ulonglong), ("max_iterations", C.c_uint)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchParams(C.Structure):
    _fields_ = [("b", c_int), ("nev", C.c_uint), ("ncv", C.c_uint), ("max_restarts", C.c_uint), ("tol", C.c_double), ("order", c_int),
                ("cholqr_passes", c_int), ("tune", c_int), ("use_graph", c_int), ("reorth", c_int)]


class BatchStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("host_seconds", C.c_double), ("setup_seconds", C.c_double), ("cycles", C.c_uint),
                ("graph_launches", C.c_uint), ("ntiles", c_int), ("format", c_int), ("rows_per_thread", c_int), ("tuned_us", C.c_double),
                ("extend_seconds", C.c_double), ("rotate_seconds", C.c_double), ("retiles", C.c_uint)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


SOLVER_API = [
    ("trl_hip_solve_csr", c_int, [c_void_p, c_int, c_size, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(SolverParams), c_void_p, c_void_p,
                                  c_void_p, c_void_p, C.POINTER(SolverResult)]),
    ("trl_hip_solve_csr_mixed", c_int, [c_void_p, c_size, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        C.POINTER(SolverParams), C.POINTER(MixedParams), c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(SolverResult),
                                        C.POINTER(SolverResult)]),
    ("trl_hip_solve_csr_shift_invert", c_int, [c_void_p, c_size, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.c_double, C.c_double,
                                               C.POINTER(SolverParams), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(SolverResult),
                                               C.POINTER(InnerStats)]),
    ("trl_hip_dense_syev", c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_dense_syev_select", c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    ("trl_hip_dense_syev_arrow_band", c_int, [c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_batch_solve_csr", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, C.POINTER(BatchParams), c_void_p, c_void_p,
                                        c_void_p, c_void_p, C.POINTER(BatchStats)]),
    ("trl_hip_mm_read", c_int, [C.c_char_p, C.POINTER(c_void_p)]),
    ("trl_hip_mm_info", c_int, [c_void_p, C.POINTER(c_size), C.POINTER(c_size), C.POINTER(c_size), c_int_p]),
    ("trl_hip_mm_copy", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("trl_hip_mm_free", None, [c_void_p]),
    ("trl_hip_mm_last_error", C.c_char_p, []),
]

_solver = None


def solver() -> C.CDLL:
    global _solver
    if _solver is None:
        hip()
        lib = _load("libtrl_hip_solver.so")
        for name, res, args in SOLVER_API:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _solver = lib
    return _solver
