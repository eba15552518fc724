"""ctypes binding of libgtpeec.so (include/gtpeec.h). The library is loaded from this package's directory; a
missing library is an error (there is no fallback path)."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .problem import DeviceError, SolverError

LIB_PATH = Path(__file__).resolve().parent / "libgtpeec.so"

GT_OK, GT_ERR_ARG, GT_ERR_OOM, GT_ERR_DEVICE, GT_ERR_NOCONV, GT_ERR_NUMERIC, GT_ERR_UNSUPPORTED = range(7)

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
i8p = C.POINTER(C.c_int8)
f64p = C.POINTER(C.c_double)


class StackDesc(C.Structure):
    _fields_ = [("n_sheets", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nxb", C.c_int32),
                ("nyb", C.c_int32), ("nbr", C.c_int32), ("n_keep", C.c_int32), ("x_sij", i32p), ("y_sij", i32p),
                ("a_keep", i32p), ("b_keep", i32p), ("node_sheet", i32p), ("px", C.c_int32 * 2),
                ("py", C.c_int32 * 2), ("kx", f64p), ("ky", f64p)]


class OpDesc(C.Structure):
    _fields_ = [("w", C.c_double), ("r_sheet", f64p), ("z_tail", f64p), ("dinv_sheet", f64p), ("dinv_tail", f64p),
                ("s_scale", C.c_double * 2), ("k", C.c_int32), ("q", f64p), ("s_in", f64p), ("s_out", f64p)]


class KernelDesc(C.Structure):
    _fields_ = [("d", C.c_double), ("mu", C.c_double), ("dz", f64p), ("near_field", f64p)]


class Segment(C.Structure):
    _fields_ = [("a", C.c_double * 3), ("b", C.c_double * 3), ("width", C.c_double), ("current", C.c_double)]


class GmresParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("bnorm", C.c_double), ("restart", C.c_int32), ("max_cycles", C.c_int32),
                ("extra_cycles", C.c_int32), ("max_its", C.c_int32), ("check_misses", C.c_int32)]


class GmresResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("products", C.c_int32),
                ("residual", C.c_double)]


# every symbol the header declares, with its ctypes signature
SIGNATURES = {
    "gt_last_error": (C.c_char_p, []),
    "gt_abi_version": (C.c_int, []),
    "gt_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gt_ctx_destroy": (None, [C.c_void_p]),
    "gt_ctx_sync": (C.c_int, [C.c_void_p]),
    "gt_ctx_trim": (C.c_int, [C.c_void_p]),
    "gt_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "gt_stack_create": (C.c_int, [C.c_void_p, C.POINTER(StackDesc), C.POINTER(C.c_void_p)]),
    "gt_stack_create_gen": (C.c_int, [C.c_void_p, C.POINTER(StackDesc), C.POINTER(KernelDesc), C.POINTER(C.c_void_p)]),
    "gt_stack_kernel": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "gt_stack_set_narrow": (C.c_int, [C.c_void_p, f64p, f64p, C.c_int64, i32p, i32p, f64p]),
    "gt_stack_lp_autotune": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p]),
    "gt_stack_set_lp_plan": (C.c_int, [C.c_void_p, i32p, i32p]),
    "gt_node_field": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
    "gt_loop_terms": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, i32p, f64p, C.c_double,
                                f64p, f64p]),
    "gt_via_coupling": (C.c_int, [C.c_void_p, C.c_int32, f64p, C.c_int32, C.POINTER(Segment), C.c_double, f64p]),
    "gt_raster_label": (C.c_int, [C.c_void_p, C.c_int32, i64p, f64p, i32p, C.c_int32, f64p, C.c_int32, f64p, i32p]),
    "gt_raster_fine": (C.c_int, [C.c_void_p, C.c_int32, i64p, f64p, i32p, C.c_int32, C.c_int32, C.c_int32, f64p,
                                 f64p, i32p, f64p, f64p, f64p, f64p, f64p, i32p]),
    "gt_raster_state": (C.c_int, [C.c_void_p, C.c_int64, f64p, i32p, C.c_int32, f64p, C.c_int32, f64p, i8p]),
    "gt_raster_cover": (C.c_int, [C.c_void_p, C.c_int64, f64p, C.c_double, C.c_double, C.c_int32, C.c_double,
                                  C.c_double, C.c_int32, f64p]),
    "gt_raster_points": (C.c_int, [C.c_void_p, C.c_int64, f64p, C.c_int64, f64p, i8p, i8p, f64p]),
    "gt_strip_coupling": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int32, C.POINTER(Segment), f64p, f64p,
                                    C.c_void_p]),
    "gt_stack_destroy": (None, [C.c_void_p]),
    "gt_stack_device_bytes": (C.c_int64, [C.c_void_p]),
    "gt_lp_matvec": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gt_lp_matvec_batch": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "gt_lp_matvec_batch_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p),
                                             C.POINTER(C.c_void_p)]),
    "gt_schur_factor": (C.c_int, [C.c_void_p, i32p, f64p]),
    "gt_stack_derive_bytes": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64)]),
    "gt_stack_create_like": (C.c_int, [C.c_void_p, C.POINTER(StackDesc), C.c_void_p, C.POINTER(C.c_void_p)]),
    "gt_matvec_on": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gt_schur_factor_nodes": (C.c_int, [C.c_void_p, i32p, i32p, i32p, f64p]),
    "gt_stack_derive": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, f64p, C.POINTER(C.c_void_p)]),
    "gt_ctx_lp_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "gt_ctx_schur_device_fronts": (C.c_int, [C.c_void_p, C.c_int32]),
    "gt_ctx_schur_leaf": (C.c_int, [C.c_void_p, C.c_int32]),
    "gt_ctx_woodbury": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "gt_ctx_schur_ordering": (C.c_int, [C.c_void_p, C.c_int32]),
    "gt_stack_woodbury": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "gt_ctx_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "gt_ctx_orthogonalization": (C.c_int, [C.c_void_p, C.c_int]),
    "gt_ctx_lp_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "gt_schur_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), i32p, i32p, i32p]),
    "gt_schur_top_stats": (C.c_int, [C.c_void_p, i32p, i32p, C.POINTER(C.c_int64)]),
    "gt_schur_solve_host": (C.c_int, [C.c_void_p, f64p, f64p]),
    "gt_schur_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gt_schur_host_factor": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, i32p, C.POINTER(C.c_void_p)]),
    "gt_schur_host_solve": (C.c_int, [C.c_void_p, f64p, f64p, C.c_int32]),
    "gt_schur_host_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), i32p, i32p, i32p]),
    "gt_schur_host_free": (None, [C.c_void_p]),
    "gt_op_create": (C.c_int, [C.c_void_p, C.POINTER(OpDesc), C.POINTER(C.c_void_p)]),
    "gt_op_destroy": (None, [C.c_void_p]),
    "gt_matvec": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gt_precond": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gt_gmres": (C.c_int, [C.c_void_p, f64p, f64p, C.POINTER(GmresParams), C.POINTER(GmresResult)]),
    "gt_gmres_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(GmresParams), C.POINTER(GmresResult)]),
    "gt_gmres_batch": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p), C.POINTER(GmresParams), C.POINTER(GmresResult)]),
    "gt_gmres_batch_fp32_basis": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                            C.POINTER(C.c_void_p), C.POINTER(GmresParams), C.POINTER(GmresResult)]),
    "gt_gmres_batch_mixed": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_void_p), C.POINTER(GmresParams), C.c_double, C.c_int32,
                                       C.POINTER(GmresResult)]),
    "gt_session_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "gt_session_destroy": (None, [C.c_void_p]),
    "gt_session_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(GmresParams), i32p]),
    "gt_session_step": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p]),
    "gt_session_poll": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(GmresResult), i32p]),
    "gt_session_device_bytes": (C.c_int64, [C.c_void_p]),
    "gt_session_plan_bytes": (C.c_int64, [C.c_void_p, C.c_int32, C.c_int32]),
    "gt_session_open": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "gt_session_open_bytes": (C.c_int64, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
}

_LIB = None


def lib():
    """The loaded library (DeviceError when it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise DeviceError(f"{LIB_PATH} is missing: build it first (python -m groundtruth_amd.build)")
        # torch first: its wheel carries its own HIP and HSA runtimes under the same soname as ROCm's. Loaded
        # after this library, torch would bring a second runtime into the process, which then finds no device;
        # loaded before it, this library binds to torch's, and the two share one runtime and device memory.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(code: int, what: str = ""):
    """Raise the exception a status code stands for (SolverError, MemoryError, DeviceError, ValueError)."""
    if code == GT_OK:
        return
    msg = lib().gt_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if code in (GT_ERR_NOCONV, GT_ERR_NUMERIC):
        raise SolverError(text)
    if code == GT_ERR_OOM:
        raise MemoryError(text)
    if code == GT_ERR_ARG:
        raise ValueError(text)
    if code == GT_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise DeviceError(text)


def ptr(a: np.ndarray, kind=f64p):
    """A ctypes pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(kind)


def cplx(a) -> np.ndarray:
    """A complex128 array as interleaved doubles (contiguous)."""
    return np.ascontiguousarray(np.asarray(a, np.complex128)).view(np.float64)


class Context:
    """One HIP device and its stream (the library's gt_ctx)."""

    _by_device: dict = {}

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().gt_ctx_create(int(device), C.byref(h)), "gt_ctx_create")
        self.handle = h
        self.device = device

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        ctx = cls._by_device.get(device)
        if ctx is None:
            ctx = cls._by_device[device] = Context(device)
        return ctx

    def sync(self):
        check(lib().gt_ctx_sync(self.handle), "gt_ctx_sync")

    def __del__(self):
        try:
            if self.handle and _LIB is not None:
                _LIB.gt_ctx_destroy(self.handle)
        except Exception:       # noqa: BLE001 - interpreter shutdown
            pass


def exported_symbols() -> list:
    """Names of the header's entry points that the library exports (every one, once built)."""
    L = lib()
    return [n for n in SIGNATURES if hasattr(L, n)]


def device_available() -> bool:
    """Whether a HIP device is visible (without touching the GPU when HIP_VISIBLE_DEVICES hides them all)."""
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "":
        return False
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:       # noqa: BLE001
        return False
