"""ctypes binding of libwesthip.so (the C ABI declared in include/west_hip.h).

The product path has no CPU fallback: if the library is not built, or no
device is usable, every call raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libwesthip.so"

_lib = None


class WestHipError(RuntimeError):
    pass


class DavidsonParams(C.Structure):
    _fields_ = [("n_pdep_eigen", C.c_int), ("n_pdep_times", C.c_int), ("n_pdep_maxiter", C.c_int),
                ("trev_pdep", C.c_double), ("trev_pdep_rel", C.c_double), ("tr2_dfpt", C.c_double)]


class DavidsonStep(C.Structure):
    _fields_ = [("dav_iter", C.c_int), ("notcnv", C.c_int), ("dfpt_dim", C.c_int), ("diago_dim", C.c_int),
                ("sternop_ncalls", C.c_long), ("dfpt_tr2", C.c_double), ("time_elap", C.c_double),
                ("time_iter", C.c_double)]


P = C.c_void_p
I = C.c_int
D = C.c_double

SIGNATURES = {
    "wh_init": ([I], I),
    "wh_finalize": ([], I),
    "wh_last_error": ([], C.c_char_p),
    "wh_system_create": ([P], I),
    "wh_system_select": ([I], I),
    "wh_system_destroy": ([I], I),
    "wh_set_grid": ([I, I, I, I, I, I, P, P, D, D, P], I),
    "wh_set_potential": ([P], I),
    "wh_set_nonlocal": ([I, P, P], I),
    "wh_set_states": ([I, P, P, P, D, D], I),
    "wh_set_coulomb": ([P], I),
    "wh_set_params": ([I, I], I),
    "wh_set_davidson_restart": ([C.c_char_p, I, I], I),
    "wh_set_precision": ([I, D], I),
    "wh_apply_sternheimerop_to_m_wfcs": ([I, P, P, P, P, I], I),
    "wh_apply_sternheimerop_to_m_wfcs_noalloc": ([I, P, P, P, P, I], I),
    "wh_linsolve_sternheimer_m_wfcts": ([I, I, P, P, P, P, D, P], I),
    "wh_dfpt": ([I, P, P, D, I, P], I),
    "wh_apply_operator": ([I, P, P, D, I], I),
    "wh_davidson_diago": ([P, P, P, P, P, I, P, P, P], I),
    "wh_solve_deflated_lanczos_w_full_ortho": ([I, I, I, P, P, P, P, P], I),
    "wh_glbrak_gamma": ([I, P, I, P, P], I),
    "wh_set_kgrid": ([I, I, I, D, P, I, P, I, I, I, I, I, P, D], I),
    "wh_set_kpoint": ([I, I, P, P, P, P, P], I),
    "wh_set_qpoint": ([I, I, P, P, P, P, P, P], I),
    "wh_set_current_k": ([I], I),
    "wh_davidson_diago_k": ([I, P, P, P, P, P, I, P, P, P], I),
    "wh_get_counters": ([P, P], I),
    "wh_get_operator_columns": ([P], I),
    "wh_get_precision_counters": ([P, P, P], I),
    "wh_profile": ([I], I),
    "wh_profile_query": ([C.c_char_p, P, P], I),
}


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise WestHipError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    # PyTorch ships its own HIP runtime with the same soname as the system one;
    # loading torch first makes the library bind to that single runtime, so
    # device tensors and library allocations live in one context.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    for name, (args, res) in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def check(rc: int, allow_positive: bool = False) -> int:
    if rc < 0 or (rc > 0 and not allow_positive):
        msg = load().wh_last_error().decode(errors="replace")
        raise WestHipError(f"libwesthip error {rc}: {msg}")
    return rc


def ptr(a) -> int:
    """Address of a numpy array (host) or a torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays passed to libwesthip must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensors passed to libwesthip must be contiguous")
        return a.data_ptr()
    raise TypeError(f"unsupported array type {type(a)}")
