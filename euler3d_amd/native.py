"""ctypes binding of libe3d_hip.so (the C ABI declared in include/e3d.h).

The library is loaded from euler3d_amd/lib/ only; there is no fallback implementation, so every
solver call either runs the HIP path or raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# E3D_LIB selects another build of the same library (the sanitized host build of the tests)
LIB_PATH = Path(os.environ["E3D_LIB"]).resolve() if os.environ.get("E3D_LIB") else \
    Path(__file__).resolve().parent / "lib" / "libe3d_hip.so"

E3D_OK = 0
ERRORS = {1: "argument", 2: "io", 3: "mesh", 4: "device", 5: "unsupported", 6: "busy", 7: "numeric"}
E3D_ERR_ARG = 1
E3D_ERR_BUSY = 6
E3D_ERR_NUMERIC = 7
# done codes of a batched case (E3D_CASE_* in include/e3d.h)
CASE_OUTCOMES = {0: "running", 1: "converged", 2: "capped", 3: "failed", 4: "stalled"}


class E3DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"e3d error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("mach", C.c_double), ("aoa_deg", C.c_double), ("gamma", C.c_double), ("cfl", C.c_double),
        ("sref", C.c_double), ("cref", C.c_double), ("ref_point", C.c_double * 3),
        ("orient_cl", C.c_int * 2), ("orient_cd", C.c_int * 2), ("orient_cm", C.c_int * 2),
        ("scheme", C.c_int), ("time_integration", C.c_int), ("residual_smoothing", C.c_int),
        ("max_iter", C.c_int), ("sampling", C.c_int),
        ("min_residual", C.c_double), ("min_aerocoeff_error", C.c_double),
    ]


class MeshInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("n_node", "n_elem", "n_face", "n_face_int", "n_ghost", "n_wall", "n_far", "n_sym")]


class PartitionInfo(C.Structure):
    _fields_ = [(k, C.c_int) for k in
                ("elem0", "n_elem", "node0", "n_node", "face_int0", "n_face_int", "ghost0", "n_ghost", "wall0",
                 "n_wall")]


class BatchOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_double), ("max_cases", C.c_int), ("wg_per_cu", C.c_int), ("xcds", C.c_int),
                ("nontemporal", C.c_int), ("use_graph", C.c_int)]


class CaseOpts(C.Structure):
    _fields_ = [("drop", C.c_double), ("max_iter", C.c_int), ("hist_cap", C.c_int), ("stall_window", C.c_int),
                ("stall_ratio", C.c_double), ("stall_level", C.c_double)]


_lib = None

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

_SIGS = {
    "e3d_last_error": (C.c_char_p, []),
    "e3d_version": (C.c_char_p, []),
    "e3d_device_count": (C.c_int, [IP]),
    "e3d_math_selfcheck": (C.c_int, [C.c_int, C.c_longlong, C.c_ulonglong, C.POINTER(C.c_ulonglong)]),
    "e3d_copy_probe": (C.c_int, [C.c_int, C.c_longlong, DP]),
    "e3d_params_default": (C.c_int, [C.POINTER(Params)]),
    "e3d_config_read": (C.c_int, [C.c_char_p, C.POINTER(Params), C.c_char_p, C.c_int, C.c_char_p, C.c_int]),
    "e3d_config_partition": (C.c_int, [C.c_char_p, C.c_int, C.c_char_p, C.c_int, IP]),
    "e3d_config_initial_mesh": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int]),
    "e3d_mesh_read": (C.c_int, [C.c_char_p, C.c_int, PP]),
    "e3d_mesh_read_partitions": (C.c_int, [C.c_int, C.POINTER(C.c_char_p), PP]),
    "e3d_mesh_partition_count": (C.c_int, [P, IP]),
    "e3d_mesh_partition_info": (C.c_int, [P, C.c_int, C.POINTER(PartitionInfo)]),
    "e3d_mesh_partition_assign": (C.c_int, [P, C.c_int, IP]),
    "e3d_mesh_partition_write": (C.c_int, [P, C.c_int, IP, C.POINTER(C.c_char_p), C.c_int]),
    "e3d_mesh_write": (C.c_int, [P, C.c_char_p, C.c_int, C.c_int]),
    "e3d_mesh_gen_box": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_char_p, PP]),
    "e3d_mesh_gen_naca": (C.c_int, [C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [PP]),
    "e3d_mesh_gen_bump": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_ulonglong, C.c_int, PP]),
    "e3d_mesh_gen_wing": (C.c_int, [C.c_int, C.c_int, C.c_int] + [C.c_double] * 5 + [PP]),
    "e3d_mesh_preprocess": (C.c_int, [P, C.c_double]),
    "e3d_mesh_get_info": (C.c_int, [P, C.POINTER(MeshInfo)]),
    "e3d_mesh_get_i32": (C.c_int, [P, C.c_char_p, IP, C.c_longlong, C.POINTER(C.c_longlong)]),
    "e3d_mesh_get_f64": (C.c_int, [P, C.c_char_p, DP, C.c_longlong, C.POINTER(C.c_longlong)]),
    "e3d_mesh_free": (None, [P]),
    "e3d_create": (C.c_int, [P, C.POINTER(Params), C.c_int, PP]),
    "e3d_set_state": (C.c_int, [P, C.POINTER(DP)]),
    "e3d_get_state": (C.c_int, [P, C.POINTER(DP)]),
    "e3d_update_bc": (C.c_int, [P]),
    "e3d_face_flux": (C.c_int, [P, DP, DP]),
    "e3d_residual": (C.c_int, [P, DP]),
    "e3d_timestep": (C.c_int, [P, DP]),
    "e3d_iterate": (C.c_int, [P, C.c_int, DP]),
    "e3d_iterate_noalloc": (C.c_int, [P, C.c_int, DP]),
    "e3d_aero": (C.c_int, [P, DP, DP]),
    "e3d_wall_cp": (C.c_int, [P, DP, C.c_longlong, C.POINTER(C.c_longlong)]),
    "e3d_run": (C.c_int, [P, DP, C.c_int, IP, IP, DP, DP, DP]),
    "e3d_last_elapsed_ms": (C.c_int, [P, DP]),
    "e3d_checkpoint_write": (C.c_int, [P, C.c_char_p]),
    "e3d_checkpoint_read": (C.c_int, [P, C.c_char_p]),
    "e3d_set_kernel_timing": (C.c_int, [P, C.c_int]),
    "e3d_kernel_timing": (C.c_int, [P, DP, IP, C.c_char_p, C.c_int]),
    "e3d_destroy": (None, [P]),
    "e3d_batch_opts_default": (C.c_int, [C.POINTER(BatchOpts)]),
    "e3d_batch_create": (C.c_int, [C.c_int, C.POINTER(BatchOpts), PP]),
    "e3d_batch_case_bytes": (C.c_int, [P, C.POINTER(CaseOpts), DP]),
    "e3d_batch_add": (C.c_int, [P, P, C.POINTER(Params), C.POINTER(CaseOpts), IP]),
    "e3d_batch_run": (C.c_int, [P, C.c_int]),
    "e3d_batch_wait": (C.c_int, [P]),
    "e3d_batch_status": (C.c_int, [P, C.c_int, IP, IP, DP, DP]),
    "e3d_batch_poll": (C.c_int, [P, IP, IP, IP, DP, C.c_int, IP]),
    "e3d_batch_result": (C.c_int, [P, C.c_int, DP, C.c_int, IP, DP, DP, DP]),
    "e3d_batch_get_state": (C.c_int, [P, C.c_int, C.POINTER(DP)]),
    "e3d_batch_remove": (C.c_int, [P, C.c_int]),
    "e3d_batch_info": (C.c_int, [P, IP, DP, DP]),
    "e3d_batch_set_tuning": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
    "e3d_batch_set_timing": (C.c_int, [P, C.c_int]),
    "e3d_batch_timing": (C.c_int, [P, DP, IP, IP]),
    "e3d_batch_destroy": (None, [P]),
}

EXPORTED = tuple(_SIGS)


def lib() -> C.CDLL:
    """Load libe3d_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise E3DError(4, f"{LIB_PATH} is missing: build it with `python -m euler3d_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != E3D_OK:
        raise E3DError(rc, lib().e3d_last_error().decode())


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(DP)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(IP)


def version() -> str:
    return lib().e3d_version().decode()


def default_params(**kw) -> Params:
    p = Params()
    check(lib().e3d_params_default(C.byref(p)))
    for k, v in kw.items():
        if k in ("ref_point", "orient_cl", "orient_cd", "orient_cm"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def has_device() -> bool:
    """True when the library sees a HIP device."""
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "":
        return False
    n = C.c_int(0)
    check(lib().e3d_device_count(C.byref(n)))
    return n.value > 0