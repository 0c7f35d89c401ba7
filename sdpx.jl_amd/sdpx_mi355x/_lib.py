"""ctypes binding of ``libsdpx_mi355x.so`` (declared in include/sdpx_mi355x.h).

The library is the only compute path: importing a function that needs it
raises when the shared object is missing, there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsdpx_mi355x.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_h = C.c_void_p

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "sdpx_last_error": [],
    "sdpx_version": [],
    "sdpx_device_count": [_ip],
    "sdpx_dense_create": [C.c_int, C.c_int, _ip, C.c_int, C.POINTER(_h)],
    "sdpx_dense_destroy": [_h],
    "sdpx_dense_memory": [_h, C.POINTER(C.c_longlong)],
    "sdpx_dense_set_block": [_h, C.c_int, _dp, _dp],
    "sdpx_dense_set_block_sparse": [_h, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _dp, _dp],
    "sdpx_dense_set_arrow": [_h, _ip],
    "sdpx_la_create": [C.c_int, C.POINTER(_h)],
    "sdpx_la_destroy": [_h],
    "sdpx_la_chol": [_h, C.c_int, _dp, C.c_longlong],
    "sdpx_la_trsm": [_h, C.c_int, C.c_int, _dp, C.c_longlong, _dp, C.c_longlong],
    "sdpx_la_trsv": [_h, C.c_int, _dp, C.c_longlong, _dp, C.c_int],
    "sdpx_la_syrk": [_h, C.c_int, C.c_int, C.c_double, _dp, C.c_longlong, C.c_double, _dp, C.c_longlong],
    "sdpx_la_gemm": [_h, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_longlong, _dp,
                     C.c_longlong, C.c_double, _dp, C.c_longlong],
    "sdpx_la_axpby": [_h, C.c_longlong, C.c_double, _dp, C.c_double, _dp],
    "sdpx_la_dot": [_h, C.c_longlong, _dp, _dp, _dp],
    "sdpx_la_norminf": [_h, C.c_longlong, _dp, _dp],
    "sdpx_dense_factor_condition": [_h, C.c_int, _dp],
    "sdpx_dense_get_arrow": [_h, _ip, _ip, _ip],
    "sdpx_dense_set_objective": [_h, _dp],
    "sdpx_dense_set_iterate": [_h, _dp, _dp, _dp],
    "sdpx_dense_get_iterate": [_h, _dp, _dp, _dp],
    "sdpx_dense_get_direction": [_h, _dp, _dp, _dp],
    "sdpx_dense_get_schur": [_h, _dp],
    "sdpx_dense_set_schur": [_h, _dp],
    "sdpx_dense_set_rhs": [_h, _dp],
    "sdpx_dense_get_rhs": [_h, _dp],
    "sdpx_dense_set_dx": [_h, _dp],
    "sdpx_dense_residuals": [_h, _dp, C.c_int, C.c_int, _dp],
    "sdpx_dense_schur_build": [_h],
    "sdpx_dense_schur_factor": [_h, _ip],
    "sdpx_dense_predictor_rhs": [_h],
    "sdpx_dense_corrector_rhs": [_h, C.c_int, C.c_double, C.c_double, C.c_int, _dp, C.c_double],
    "sdpx_dense_solve": [_h],
    "sdpx_dense_direction_residual": [_h, _dp],
    "sdpx_dense_refine": [_h, C.c_double, C.c_int, C.c_int, _ip, _dp],
    "sdpx_dense_direction_blocks": [_h],
    "sdpx_dense_complementarity": [_h, C.c_double, C.c_double, _dp],
    "sdpx_dense_fraction_to_boundary": [_h, C.c_double, C.c_double, _dp, _dp],
    "sdpx_dense_backtrack": [_h, C.c_double, C.c_double, C.c_double, _dp, _dp],
    "sdpx_dense_update": [_h, C.c_double, C.c_double, _dp],
    "sdpx_dense_accepted_residuals": [_h, C.c_double, C.c_double, C.c_double, _dp],
    "sdpx_dense_objectives": [_h, _dp, _dp],
    "sdpx_dense_solution_residuals": [_h, _dp, _dp],
    "sdpx_dense_scale_blocks": [_h, C.c_int, C.c_double],
    "sdpx_dense_factor_quality": [_h, _dp],
    "sdpx_dense_synchronize": [_h],
    "sdpx_dense_set_profiling": [_h, C.c_int],
    "sdpx_dense_kernel_times": [_h, C.c_int, _dp],
    "sdpx_dense_schur_split": [_h, _ip],
    "sdpx_dense_autotune": [_h, _dp],
    "sdpx_dense_set_tiling": [_h, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "sdpx_gram_work_list": [C.c_int, C.c_int, C.c_int, _ip],
    "sdpx_dense_set_precision": [_h, C.c_int],
    "sdpx_dense_set_factor_precision": [_h, C.c_int],
    "sdpx_dense_set_panel_precision": [_h, C.c_int],
    "sdpx_dense_equilibrate": [_h, C.c_int, _dp, _dp, _dp, _ip],
    "sdpx_dense_set_equalities": [_h, C.c_int, _dp],
    "sdpx_dense_equality_factor": [_h, C.c_double, _dp],
    "sdpx_dense_solve_equality": [_h, _dp, _dp],
    "sdpx_dense_get_equality": [_h, _dp, _dp, _dp],
    "sdpx_dense_get_data": [_h, _dp, _dp],
    "sdpx_dense_cold_start": [_h, C.c_double, C.c_double, C.c_double, _dp, _dp],
    "sdpx_dense_exact_residual": [_h, _dp],
    "sdpx_dense_refine_exact": [_h, C.c_double, C.c_int, _ip, _dp],
    "sdpx_batch_default_options": [C.c_void_p, C.c_double],
    "sdpx_batch_create": [C.c_int, C.c_int, _ip, _ip, C.POINTER(_h)],
    "sdpx_batch_destroy": [_h],
    "sdpx_batch_memory": [_h, C.POINTER(C.c_longlong)],
    "sdpx_batch_set_dense": [_h, _dp, _dp, _dp],
    "sdpx_batch_set_coo": [_h, C.POINTER(C.c_longlong), _ip, _ip, _dp, _dp, _dp],
    "sdpx_batch_set_block_coo": [_h, _ip, _ip, _ip, _ip, _dp, _dp, _dp],
    "sdpx_block_coo_convert": [C.c_int, C.c_int, _ip, _ip, _ip, _ip, _dp, C.POINTER(C.c_longlong), _ip, _ip, _dp,
                               C.POINTER(C.c_longlong)],
    "sdpx_batch_set_start": [_h, _dp, _dp, _dp],
    "sdpx_batch_solve": [_h, C.c_void_p, _ip],
    "sdpx_batch_get_result": [_h, _ip, _ip, _dp, _dp, _dp, _dp, _dp],
    "sdpx_batch_get_schur": [_h, C.c_int, _dp],
    "sdpx_batch_get_trace": [_h, C.c_int, _dp, C.c_int, _ip],
    "sdpx_batch_adjoint": [_h, _dp, _dp, _dp, _dp],
    "sdpx_batch_center": [_h, C.c_int],
    "sdpx_batch_center_at": [_h, C.c_int, _dp],
    "sdpx_batch_autotune": [_h, C.c_int, _dp],
    "sdpx_batch_set_gram_order": [_h, C.c_int],
    "sdpx_batch_restore_checkpoint": [_h, _ip, _dp],
    "sdpx_batch_schur_at_start": [_h],
    "sdpx_batch_set_precision": [_h, C.c_int],
    "sdpx_batch_get_mixed_iterations": [_h, _ip],
    "sdpx_batch_get_reasons": [_h, _ip],
    "sdpx_batch_policy_select": [C.c_void_p, _dp, C.c_int, _dp, C.c_int, _dp],
    "sdpx_batch_policy_record": [C.c_void_p, _dp, _dp],
    "sdpx_soc_create": [C.c_int, C.c_int, _ip, C.POINTER(_h)],
    "sdpx_soc_destroy": [_h],
    "sdpx_soc_memory": [_h, C.POINTER(C.c_longlong)],
    "sdpx_soc_set_vector": [_h, C.c_int, _dp],
    "sdpx_soc_get_vector": [_h, C.c_int, _dp],
    "sdpx_soc_copy_vector": [_h, C.c_int, C.c_int],
    "sdpx_soc_get_scaling": [_h, _dp, _dp, _ip],
    "sdpx_soc_nt_scaling": [_h, C.c_int, C.c_int, C.c_int, C.c_int, _ip],
    "sdpx_soc_apply": [_h, C.c_int, C.c_int, C.c_int, C.c_int],
    "sdpx_soc_jordan": [_h, C.c_int, C.c_int, C.c_int],
    "sdpx_soc_jordan_solve": [_h, C.c_int, C.c_int, C.c_int, _ip],
    "sdpx_soc_inverse": [_h, C.c_int, C.c_int, _ip],
    "sdpx_soc_fraction_to_boundary": [_h, C.c_int, C.c_int, _dp, _dp],
    "sdpx_soc_set_constraints": [_h, C.c_int, _dp],
    "sdpx_soc_metric": [_h, C.c_int, _dp, _dp, _ip],
    "sdpx_soc_fraction_to_boundary_noalloc": [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p],
    "sdpx_soc_margins": [_h, C.c_int, _dp, _ip],
    "sdpx_soc_strict_step": [_h, C.c_int, C.c_int, C.c_double, C.c_int, _dp],
    "sdpx_soc_hkm_metric": [_h, C.c_int, C.c_int, C.c_int, _ip],
    "sdpx_soc_hkm_contract": [_h, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                              C.c_int],
    "sdpx_soc_set_profiling": [_h, C.c_int],
    "sdpx_soc_last_kernel_ms": [_h, _dp],
    "sdpx_soc_synchronize": [_h],
}
_RESTYPES = {"sdpx_last_error": C.c_char_p}

_LIB = None


class SDPXError(RuntimeError):
    """A C-ABI call failed (negative status)."""


def load() -> C.CDLL:
    """Load the HIP library; raises when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise SDPXError(
            f"{LIB_PATH.name} is not built (run __graft_entry__.build()); "
            "the MI355X provider has no CPU fallback"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _LIB = lib
    return lib


def check(status: int, what: str) -> int:
    """Raise on negative status; return the non-negative status (0 ok, 1 not PD, 2 non-finite)."""
    if status < 0:
        msg = load().sdpx_last_error()
        raise SDPXError(f"{what} failed ({status}): {msg.decode() if msg else ''}")
    return status


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None
