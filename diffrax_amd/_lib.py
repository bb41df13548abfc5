"""ctypes binding of libdiffrax_amd.so (the C ABI in include/diffrax_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
diffrax_amd/csrc``).  There is no fallback: if the shared object is missing or a GPU
is not available, every solve raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiffrax_amd.so")

# enums (include/diffrax_amd.h)
TSIT5, DOPRI5, HEUN = 0, 1, 2
PID, CONSTANT = 0, 1
RESUME_SOLVER, RESUME_CONTROLLER, RESUME_MADE_JUMP = 1, 2, 4
VF_LV, VF_VDP, VF_LINEAR, VF_MLP, VF_OU = 0, 1, 2, 3, 4
F64, F32S_F64T, F32, BF16_F32 = 0, 1, 2, 3
SUCCESSFUL, MAX_STEPS_REACHED, DT_MIN_REACHED = 0, 1, 2


class ErkProblem(ctypes.Structure):
    _fields_ = [
        ("solver", ctypes.c_int32), ("controller", ctypes.c_int32),
        ("vf", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("d", ctypes.c_int32), ("hidden", ctypes.c_int32), ("n_ts", ctypes.c_int32),
        ("save_t0", ctypes.c_int32), ("save_t1", ctypes.c_int32),
        ("save_steps", ctypes.c_int32), ("max_steps", ctypes.c_int32),
        ("force_dtmin", ctypes.c_int32),
        ("batch", ctypes.c_int64), ("t1_stride", ctypes.c_int64),
        ("ts_stride", ctypes.c_int64), ("param_stride", ctypes.c_int64),
        ("t0", ctypes.c_double), ("dt0", ctypes.c_double),
        ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
        ("pcoeff", ctypes.c_double), ("icoeff", ctypes.c_double), ("dcoeff", ctypes.c_double),
        ("dtmin", ctypes.c_double), ("dtmax", ctypes.c_double),
        ("factormin", ctypes.c_double), ("factormax", ctypes.c_double),
        ("safety", ctypes.c_double), ("error_order", ctypes.c_double),
        ("n_step_ts", ctypes.c_int32), ("n_jump_ts", ctypes.c_int32),
        ("n_events", ctypes.c_int32), ("event_root_find", ctypes.c_int32),
        ("ckpt_every", ctypes.c_int32), ("resume", ctypes.c_int32),
    ]


class ErkIO(ctypes.Structure):
    _fields_ = [
        ("y0", ctypes.c_void_p), ("t1", ctypes.c_void_p), ("ts", ctypes.c_void_p),
        ("params", ctypes.c_void_p), ("ys", ctypes.c_void_p), ("ts_out", ctypes.c_void_p),
        ("stats", ctypes.c_void_p), ("result", ctypes.c_void_p), ("grid", ctypes.c_void_p),
        ("dense_ts", ctypes.c_void_p), ("dense_ys", ctypes.c_void_p), ("dense_k", ctypes.c_void_p),
        ("step_ts", ctypes.c_void_p), ("jump_ts", ctypes.c_void_p),
        ("event_coeffs", ctypes.c_void_p), ("event_flags", ctypes.c_void_p),
        ("event_mask", ctypes.c_void_p), ("grid_y", ctypes.c_void_p),
        ("rtol", ctypes.c_void_p), ("atol", ctypes.c_void_p), ("ckpt", ctypes.c_void_p),
        ("state_in", ctypes.c_void_p), ("state_out", ctypes.c_void_p),
    ]


class DenseIO(ctypes.Structure):
    _fields_ = [
        ("dense_ts", ctypes.c_void_p), ("dense_ys", ctypes.c_void_p), ("dense_k", ctypes.c_void_p),
        ("stats", ctypes.c_void_p), ("t1", ctypes.c_void_p), ("tq", ctypes.c_void_p),
        ("n_q", ctypes.c_int64), ("tq_stride", ctypes.c_int64), ("left", ctypes.c_int32),
        ("derivative", ctypes.c_int32), ("out", ctypes.c_void_p),
    ]


class AdjointIO(ctypes.Structure):
    _fields_ = [
        ("ckpt", ctypes.c_void_p), ("stats", ctypes.c_void_p), ("ys_bar", ctypes.c_void_p),
        ("y0_bar", ctypes.c_void_p), ("params_bar", ctypes.c_void_p),
    ]


class BacksolveIO(ctypes.Structure):
    _fields_ = [
        ("params", ctypes.c_void_p), ("ys", ctypes.c_void_p), ("ts_out", ctypes.c_void_p),
        ("ys_bar", ctypes.c_void_p), ("y0_bar", ctypes.c_void_p),
        ("params_bar", ctypes.c_void_p), ("result", ctypes.c_void_p),
    ]


class Launch(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("lanes", ctypes.c_int32)]


# Exported symbols and their signatures: the single source of truth for the loader test.
_SIGS = {
    "dfx_version": ([], ctypes.c_int),
    "dfx_strerror": ([ctypes.c_int], ctypes.c_char_p),
    "dfx_erk_n_out": ([ctypes.POINTER(ErkProblem)], ctypes.c_int),
    "dfx_vf_param_count": ([ctypes.c_int32, ctypes.c_int32, ctypes.c_int32], ctypes.c_int64),
    "dfx_tableau": ([ctypes.c_int32] + [ctypes.POINTER(ctypes.c_double)] * 4
                    + [ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "dfx_erk_workspace_bytes": ([ctypes.POINTER(ErkProblem), ctypes.POINTER(ctypes.c_size_t)],
                                ctypes.c_int),
    "dfx_erk_solve": ([ctypes.POINTER(ErkProblem), ctypes.POINTER(ErkIO), ctypes.c_void_p,
                       ctypes.c_size_t, ctypes.POINTER(Launch), ctypes.c_void_p], ctypes.c_int),
    "dfx_erk_dense_eval": ([ctypes.POINTER(ErkProblem), ctypes.POINTER(DenseIO), ctypes.c_void_p],
                           ctypes.c_int),
    "dfx_erk_backsolve": ([ctypes.POINTER(ErkProblem), ctypes.POINTER(BacksolveIO),
                           ctypes.c_void_p], ctypes.c_int),
    "dfx_erk_ckpt_layout": ([ctypes.POINTER(ErkProblem), ctypes.POINTER(ctypes.c_int32),
                             ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "dfx_erk_adjoint_workspace_bytes": ([ctypes.POINTER(ErkProblem),
                                         ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "dfx_erk_adjoint": ([ctypes.POINTER(ErkProblem), ctypes.POINTER(ErkIO),
                         ctypes.POINTER(AdjointIO), ctypes.c_void_p, ctypes.c_size_t,
                         ctypes.c_void_p], ctypes.c_int),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libdiffrax_amd.so (raises LibraryMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is missing: build it with `make -C diffrax_amd/csrc` "
            "(or __graft_entry__.build()). diffrax_amd has no CPU fallback.")
    # torch first, so the HIP runtime torch ships is the one the library binds to.
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    for name, (argtypes, restype) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


def check(rc, what="diffrax_amd call"):
    if rc != 0:
        msg = load().dfx_strerror(rc).decode()
        raise RuntimeError(f"{what} failed: {msg} (code {rc})")
