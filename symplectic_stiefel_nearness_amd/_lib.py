"""ctypes binding of libssn.so (the C ABI declared in include/ssn.h).

Tensors cross the boundary as raw device pointers plus int64 strides; the stream is
torch's current HIP stream. No torch types appear in the C signatures.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libssn.so"

SSN_OK = 0
SSN_ERR_ARGS = -1
SSN_ERR_HIP = -2
SSN_ERR_WORKSPACE = -3
SSN_ERR_UNSUPPORTED = -4
SSN_ERR_INIT = -5
SSN_STAGE_GRAM, SSN_STAGE_DUAL, SSN_STAGE_RECOVER, SSN_STAGE_ALL = 1, 2, 4, 7   # ssn_solve_uneven_stages
SSN_STATIC_SCRATCH_BYTES = 128 << 20   # include/ssn.h: the static workspace of ssn_solve_static

ST_OK, ST_NO_DESCENT, ST_NEWTON_CAP, ST_NONFINITE, ST_EIGH = 0, 1, 2, 3, 4

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_longlong
_S = ctypes.c_size_t

class SsnTarget(ctypes.Structure):
    """Mirror of `ssn_target` (include/ssn.h)."""
    _fields_ = [("T", ctypes.c_void_p), ("sT_r", ctypes.c_longlong),
                ("X", ctypes.c_void_p), ("sX_r", ctypes.c_longlong),
                ("omega", ctypes.c_void_p), ("ldo", ctypes.c_longlong),
                ("iters", ctypes.c_void_p), ("status", ctypes.c_void_p),
                ("counters", ctypes.c_void_p),
                ("n", ctypes.c_int), ("m", ctypes.c_int),
                ("gram_chunk_rows", ctypes.c_int), ("rec_rows", ctypes.c_int)]


_TP = ctypes.POINTER(SsnTarget)


class SsnOptions(ctypes.Structure):
    """Mirror of `ssn_options` (include/ssn.h)."""
    _fields_ = [("max_newton", ctypes.c_int), ("max_cg", ctypes.c_int),
                ("max_backtrack", ctypes.c_int), ("split", ctypes.c_int),
                ("feas_tol", ctypes.c_double), ("shared_device", ctypes.c_int)]


_OP = ctypes.POINTER(SsnOptions)

# name -> (restype, argtypes); must match include/ssn.h exactly
SIGNATURES = {
    "ssn_init": (_I, []),
    "ssn_version": (ctypes.c_char_p, []),
    "ssn_padded_order": (_I, [_I]),
    "ssn_last_hip_error": (_I, []),
    "ssn_status_string": (ctypes.c_char_p, [_I]),
    "ssn_host_copy": (_I, [_P, _P, _S, _I]),
    "ssn_workspace_bytes": (_S, [_I, _I, _I]),
    "ssn_gram": (_I, [_P, _P, _L, _L, _I, _I, _I, _P, _P, _L, _I, _P, _S]),
    "ssn_dual": (_I, [_P, _P, _P, _L, _I, _I, _I, _P, _L, _I, _P, _P, _P, _P, _P, _P, _S]),
    "ssn_dual_eigh": (_I, [_P, _P, _L, _I, _I, _I, _P, _S]),
    "ssn_dual_newton": (_I, [_P, _P, _P, _L, _I, _I, _I, _P, _L, _I, _P, _P, _P, _P, _P, _P, _S]),
    "ssn_recover": (_I, [_P, _P, _L, _L, _P, _P, _I, _I, _I, _P, _L, _L, _P, _S]),
    "ssn_solve": (_I, [_P, _P, _L, _L, _I, _I, _I, _P, _L, _L, _P, _L, _I, _P, _P, _P, _P, _S]),
    "ssn_solve_static": (_I, [_P, _P, _L, _L, _I, _I, _I, _P, _L, _L, _P, _L, _I, _P, _P, _P]),
    "ssn_default_options": (None, [_OP]),
    "ssn_solve_opts": (_I, [_P, _OP, _P, _L, _L, _I, _I, _I, _P, _L, _L, _P, _L, _I, _P, _P, _P, _P, _S]),
    "ssn_backward": (_I, [_P, _P, _L, _L, _P, _L, _L, _I, _I, _I, _P, _L, _L, _P, _P, _S]),
    "ssn_gram_split": (_I, [_P, _P, _L, _L, _I, _I, _I, _P, _P, _L, _I, _P, _S]),
    "ssn_recover_split": (_I, [_P, _P, _L, _L, _P, _P, _I, _I, _I, _P, _L, _L, _P, _S]),
    "ssn_solve_split": (_I, [_P, _P, _L, _L, _I, _I, _I, _P, _L, _L, _P, _L, _I, _P, _P, _P, _P, _S]),
    "ssn_uneven_workspace_bytes": (_S, [_TP, _I]),
    "ssn_uneven_plan_bytes": (_S, [_TP, _I]),
    "ssn_uneven_plan": (_I, [_TP, _I, _P, _S, _P, _S]),
    "ssn_solve_uneven": (_I, [_P, _P, _P, _S]),
    "ssn_solve_uneven_stages": (_I, [_P, _P, _P, _S, _I]),
}

_lib = None


def load(path: Path | str | None = None):
    """Load libssn.so once and bind every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m symplectic_stiefel_nearness_amd.build` "
            "(there is no CPU fallback for the solver)")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc == SSN_OK:
        return
    lib = load()
    if rc == SSN_ERR_ARGS:
        raise AssertionError(f"{what}: invalid arguments")
    if rc == SSN_ERR_WORKSPACE:
        raise AssertionError(f"{what}: workspace too small")
    if rc == SSN_ERR_UNSUPPORTED:
        raise ValueError(f"{what}: unsupported shape for this device")
    raise RuntimeError(f"{what}: HIP error {lib.ssn_last_hip_error()} (code {rc})")
