"""ctypes binding of libbbdg.so (the C ABI declared in include/bbdg.h).

The library is built in-tree (``bernsteinbasis.jl_amd/libbbdg.so``).  There is no
fallback: if it is missing, every compute entry point raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(os.path.dirname(HERE), "libbbdg.so")

BB_OK, BB_ERR_ARG, BB_ERR_NULL, BB_ERR_HIP, BB_ERR_UNSUPPORTED, BB_ERR_STATE, BB_ERR_NOMEM, BB_ERR_NONFINITE = range(8)
BB_F64, BB_F32 = 0, 1
BB_PDE_WAVE, BB_PDE_MAXWELL = 0, 1
BB_FLUX_REFERENCE, BB_FLUX_UPWIND = 0, 1
BB_LIFT_PLAIN, BB_LIFT_ACC64, BB_LIFT_TWOSUM, BB_LIFT_ACC64_MFMA = 0, 1, 2, 3
BB_PREC_F64, BB_PREC_MIXED = 0, 1
BB_PATH_PLAIN, BB_PATH_FUSED = 0, 1

ERR_NAMES = {1: "BB_ERR_ARG", 2: "BB_ERR_NULL", 3: "BB_ERR_HIP", 4: "BB_ERR_UNSUPPORTED", 5: "BB_ERR_STATE",
             6: "BB_ERR_NOMEM", 7: "BB_ERR_NONFINITE"}


class bb_layout(C.Structure):
    _fields_ = [("elem", C.c_int64), ("node", C.c_int64), ("field", C.c_int64)]


class BBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


# exported symbol -> (restype, argtypes)
vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
L = bb_layout
SIGS = {
    "bb_version": (C.c_char_p, []),
    "bb_last_error": (C.c_char_p, [vp]),
    "bb_device_count": (i32, [C.POINTER(C.c_int)]),
    "bb_ctx_create": (i32, [i32, C.POINTER(vp)]),
    "bb_ctx_destroy": (i32, [vp]),
    "bb_ctx_stream": (i32, [vp, C.POINTER(vp)]),
    "bb_device_info": (i32, [vp, C.c_char_p, i32, C.c_char_p, i32, C.POINTER(C.c_int)]),
    "bb_malloc": (i32, [vp, C.c_size_t, C.POINTER(vp)]),
    "bb_free": (i32, [vp, vp]),
    "bb_memcpy_h2d": (i32, [vp, vp, vp, C.c_size_t]),
    "bb_memcpy_d2h": (i32, [vp, vp, vp, C.c_size_t]),
    "bb_memcpy_d2d": (i32, [vp, vp, vp, C.c_size_t]),
    "bb_fill_zero": (i32, [vp, vp, C.c_size_t]),
    "bb_stream_copy": (i32, [vp, vp, vp, C.c_size_t]),
    "bb_synchronize": (i32, [vp]),
    "bb_count_nonfinite": (i32, [vp, i32, vp, i64, C.POINTER(C.c_int64)]),
    "bb_timer_start": (i32, [vp]),
    "bb_timer_stop": (i32, [vp, C.POINTER(C.c_float)]),
    "bb_tables": (i32, [i32, vp, vp, vp]),
    "bb_linear_to_ijkl": (i32, [i32, vp]),
    "bb_linear_to_ijk_2d": (i32, [i32, vp]),
    "bb_fmask": (i32, [i32, vp]),
    "bb_deriv_entry": (i32, [i32, i32, i32, i32, C.POINTER(C.c_double)]),
    "bb_elevation_entry": (i32, [i32, i32, i32, C.POINTER(C.c_double)]),
    "bb_deriv_apply": (i32, [vp, i32, i32, i32, i64, i32, vp, L, vp, L]),
    "bb_grad_apply": (i32, [vp, i32, i32, i64, i32, vp, L, vp, vp, vp, vp, L]),
    "bb_lift_apply": (i32, [vp, i32, i32, i64, i32, i32, vp, L, vp, L]),
    "bb_face_lift_apply": (i32, [vp, i32, i32, i64, i32, i32, i32, vp, L, vp, L]),
    "bb_elevate_apply": (i32, [vp, i32, i32, i64, i32, vp, L, vp, L]),
    "bb_face_gather": (i32, [vp, i32, i32, i64, i32, vp, L, vp, L]),
    "bb_interface_flux": (i32, [vp, i32, i32, i32, i64, vp, L, vp, vp, vp, vp, vp, vp, vp, L]),
    "bb_solver_create": (i32, [vp, i32, i32, i32, i64, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]),
    "bb_solver_destroy": (i32, [vp]),
    "bb_solver_set_state": (i32, [vp, vp]),
    "bb_solver_get_state": (i32, [vp, vp]),
    "bb_solver_state_ptr": (i32, [vp, C.POINTER(vp)]),
    "bb_solver_check_state": (i32, [vp, C.POINTER(C.c_int64)]),
    "bb_solver_rhs": (i32, [vp, i32, i32, vp]),
    "bb_solver_step": (i32, [vp, i32, i32, dbl, i32]),
    "bb_solver_graph_capture": (i32, [vp, i32, i32, dbl, i32]),
    "bb_solver_graph_launch": (i32, [vp]),
    "bb_solver_autotune": (i32, [vp, i32, C.POINTER(C.c_int)]),
    "bb_solver_set_tuning": (i32, [vp, i32, i32]),
    "bb_solver_stage_bytes": (i32, [vp, i32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bb_solver_step_noalloc": (i32, [vp, i32, vp, dbl, i32]),
    "bb_gemm": (i32, [vp, i32, i64, i64, i64, dbl, vp, i64, vp, i64, dbl, vp, i64]),
    "bb_pde_flux_dirs": (i32, [vp, i32, i32, i64, vp, vp, vp, vp]),
    "bb_rhs_combine": (i32, [vp, i32, i32, i64, vp, vp, vp, vp]),
    "bb_batch_create": (i32, [vp, i32, i32, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64, C.POINTER(vp)]),
    "bb_batch_destroy": (i32, [vp]),
    "bb_batch_state_words": (i32, [vp, C.POINTER(C.c_int64)]),
    "bb_batch_set_state": (i32, [vp, vp]),
    "bb_batch_get_state": (i32, [vp, vp]),
    "bb_batch_check_state": (i32, [vp, C.POINTER(C.c_int64)]),
    "bb_batch_rhs": (i32, [vp, i32, vp]),
    "bb_batch_step": (i32, [vp, i32, dbl, i32]),
    "bb_batch_graph_capture": (i32, [vp, i32, dbl, i32]),
    "bb_batch_graph_launch": (i32, [vp]),
    "bb_batch_autotune": (i32, [vp, i32, C.POINTER(C.c_int)]),
    "bb_batch_set_tuning": (i32, [vp, i32, i32, i32]),
}

_lib = None

# Library objects still open when the interpreter exits are released by an exit hook, while
# the HIP runtime is still up: arrays, solvers and batches first, then their contexts.
# (Left to garbage collection at the very end of the process, a release could run after
# the runtime's own teardown.)
_open_objs = weakref.WeakSet()
_open_ctxs = weakref.WeakSet()


def track(obj, context=False):
    (_open_ctxs if context else _open_objs).add(obj)
    return obj


@atexit.register
def _release_all():
    for group, meth in ((_open_objs, None), (_open_ctxs, "close")):
        for o in list(group):
            try:
                getattr(o, meth or ("free" if hasattr(o, "free") else "close"))()
            except Exception:
                pass


def lib():
    """Load libbbdg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            raise RuntimeError(f"{SO} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        h = C.CDLL(SO)
        for name, (res, args) in SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(rc, ctx=None):
    if rc != BB_OK:
        msg = lib().bb_last_error(ctx).decode() if ctx else ""
        raise BBError(rc, msg)
