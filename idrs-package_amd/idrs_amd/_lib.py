"""ctypes binding of libidrs_hip.so (declarations in include/idrs_capi.h).

The library is the only compute path: importing this module fails loudly
when the shared object is missing, and every call that returns an error
status raises IdrsError with the library's message.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libidrs_hip.so")

IDRS_OK = 0
IDRS_ERR_ARG = -1
IDRS_ERR_HIP = -2
IDRS_ERR_UNSUPPORTED = -3
IDRS_ERR_INTERNAL = -4


class IdrsError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"idrs status {status}: {message}")
        self.status = status


class idrs_opts_t(ctypes.Structure):
    _fields_ = [
        ("has_M1", c_int32),
        ("has_tolerance", c_int32),
        ("has_maximum_iterations", c_int32),
        ("has_variant", c_int32),
        ("M1", c_void_p),
        ("tolerance", c_double),
        ("maximum_iterations", c_int32),
        ("variant", c_int32),
        ("x0", c_void_p),
        ("U0", c_void_p),
        ("omega", c_void_p),
        ("n_omega", c_int32),
        ("resvec_len", c_int32),
        ("resvec", c_void_p),
        ("PT_raw", c_void_p),
        ("p_seed", c_uint64),
        ("device_pointers", c_int32),
        ("time_spmv", c_int32),
        ("graph_mode", c_int32),
        ("H_nritz", c_int32),
        ("H", c_void_p),
    ]


class idrs_krylov_opts_t(ctypes.Structure):
    _fields_ = [
        ("has_M1", c_int32),
        ("has_tolerance", c_int32),
        ("has_maximum_iterations", c_int32),
        ("has_variant", c_int32),
        ("M1", c_void_p),
        ("tolerance", c_double),
        ("maximum_iterations", c_int32),
        ("variant", c_int32),
        ("has_inner_s", c_int32),
        ("inner_s", c_int32),
        ("has_inner_tolerance", c_int32),
        ("has_inner_iterations", c_int32),
        ("inner_tolerance", c_double),
        ("inner_iterations", c_int32),
        ("n_omega", c_int32),
        ("x0", c_void_p),
        ("omega", c_void_p),
        ("resvec", c_void_p),
        ("resvec_len", c_int32),
        ("device_pointers", c_int32),
        ("PT_raw", c_void_p),
        ("p_seed", c_uint64),
        ("colfac", c_void_p),
    ]


class idrs_stats_t(ctypes.Structure):
    _fields_ = [
        ("flag", c_int32),
        ("iterations", c_int32),
        ("relres", c_double),
        ("solve_ms", c_double),
        ("spmv_ms", c_double),
        ("spmv_launches", c_int32),
        ("reserved", c_int32),
    ]


# name -> (restype, argtypes)
_PROTOS = {
    "idrs_last_error": (ctypes.c_char_p, []),
    "idrs_version": (c_int, []),
    "idrs_device_count": (c_int, []),
    "idrs_crs_format_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "idrs_crs_format_c128": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "idrs_coo_format_f64": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "idrs_coo_format_c128": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "idrs_cds_format_f64": (c_int, [c_int, c_int, c_int, c_void_p, POINTER(c_void_p)]),
    "idrs_cds_format_c128": (c_int, [c_int, c_int, c_int, c_void_p, POINTER(c_void_p)]),
    "idrs_dia_format_f64": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, POINTER(c_void_p)]),
    "idrs_dia_format_c128": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, POINTER(c_void_p)]),
    "idrs_user_matrix_destroy": (c_int, [c_void_p]),
    "idrs_mm_info": (c_int, [ctypes.c_char_p, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_int),
                             POINTER(c_int), POINTER(c_int64)]),
    "idrs_mm_matrix": (c_int, [ctypes.c_char_p, c_int, POINTER(c_void_p)]),
    "idrs_mm_dense": (c_int, [ctypes.c_char_p, c_void_p]),
    "idrs_user_matrix_autotune": (c_int, [c_void_p, c_int, POINTER(c_int)]),
    "idrs_user_matrix_set_group": (c_int, [c_void_p, c_int]),
    "idrs_user_matrix_layout": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int)]),
    "idrs_user_matrix_set_row_slabs": (c_int, [c_void_p, c_int]),
    "idrs_user_matrix_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int64), POINTER(c_int)]),
    "idrs_user_mv_f64": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_user_mv_c128": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_real_diagonal": (c_int, [c_void_p, c_void_p]),
    "idrs_complex_diagonal": (c_int, [c_void_p, c_void_p]),
    "idrs_matrix_create": (c_int, [c_void_p, POINTER(c_void_p)]),
    "idrs_matrix_destroy": (c_int, [c_void_p]),
    "idrs_matrix_set_precon_f64": (c_int, [c_void_p, c_int, c_void_p]),
    "idrs_matrix_set_precon_c128": (c_int, [c_void_p, c_int, c_void_p]),
    "idrs_matrix_stored_values": (c_int, [c_void_p, POINTER(c_int64), POINTER(c_int)]),
    "idrs_matrix_set_seed_f64": (c_int, [c_void_p, c_double]),
    "idrs_matrix_set_seed_c128": (c_int, [c_void_p, c_void_p]),
    "idrs_real_preconditioner": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "idrs_complex_preconditioner": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "idrs_matrix_get_alpha": (c_int, [c_void_p, c_void_p, c_int, POINTER(c_int)]),
    "idrs_matrix_set_alpha": (c_int, [c_void_p, c_int, c_void_p]),
    "idrs_chebyshev_coeffs": (c_int, [c_int, c_void_p, c_int, c_void_p]),
    "idrs_neumann_coeffs": (c_int, [c_int, c_void_p, c_int, c_void_p]),
    "idrs_compute_alpha_coeffs": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p]),
    "idrs_matvec_f64": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_matvec_c128": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_matvec_block_f64": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "idrs_matvec_block_c128": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "idrs_precon_f64": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_precon_c128": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_scaleback_f64": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_scaleback_c128": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_idrs_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_opts_t), c_void_p, POINTER(idrs_stats_t)]),
    "idrs_idrs_c128": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_opts_t), c_void_p, POINTER(idrs_stats_t)]),
    "idrs_idrs_f64_noalloc": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_opts_t), c_void_p,
                                      POINTER(idrs_stats_t)]),
    "idrs_idrs_ritz_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_opts_t), c_void_p,
                                   POINTER(idrs_stats_t), c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int,
                                   POINTER(c_int), POINTER(c_int)]),
    "idrs_idrs_ritz_c128": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_opts_t), c_void_p,
                                    POINTER(idrs_stats_t), c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int,
                                    POINTER(c_int), POINTER(c_int)]),
    "idrs_qr_eig": (c_int, [c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, POINTER(c_int)]),
    "idrs_matrix_set_mass": (c_int, [c_void_p, c_void_p]),
    "idrs_matrix_set_damping": (c_int, [c_void_p, c_void_p]),
    "idrs_matrix_set_flags": (c_int, [c_void_p, c_int, c_int]),
    "idrs_matrix_rows": (c_int, [c_void_p, POINTER(c_int)]),
    "idrs_matrix_set_precon_diag": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "idrs_matrix_set_precon_lu": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "idrs_matrix_share_lu": (c_int, [c_void_p, c_void_p]),
    "idrs_matrix_set_shift_invert": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "idrs_ms_add_deflation_f64": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_ms_add_deflation_c128": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_make_deflation_f64": (c_int, [c_void_p, c_int, c_void_p]),
    "idrs_make_deflation_c128": (c_int, [c_void_p, c_int, c_void_p]),
    "idrs_pmul_f64": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_pmul_c128": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_system_mv_f64": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_system_mv_c128": (c_int, [c_void_p, c_void_p, c_void_p]),
    "idrs_add_deflation_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_add_deflation_c128": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_scaleback_b_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_scaleback_b_c128": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_ms_scaleback_b_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_ms_scaleback_b_c128": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "idrs_filter_ritzval": (c_int, [c_int, c_void_p, c_void_p]),
    "idrs_analyse_spectrum": (c_int, [c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "idrs_si_power": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_uint64, c_uint64, c_void_p]),
    "idrs_qmridr_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_krylov_opts_t), c_void_p,
                                POINTER(idrs_stats_t)]),
    "idrs_qmridr_c128": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_krylov_opts_t), c_void_p,
                                 POINTER(idrs_stats_t)]),
    "idrs_msidrs_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, POINTER(idrs_krylov_opts_t),
                                c_void_p, c_void_p, POINTER(idrs_stats_t)]),
    "idrs_msidrs_c128": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, POINTER(idrs_krylov_opts_t),
                                 c_void_p, c_void_p, POINTER(idrs_stats_t)]),
    "idrs_msqmridr_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, POINTER(idrs_krylov_opts_t),
                                  c_void_p, c_void_p, POINTER(idrs_stats_t)]),
    "idrs_ms_mixed_c128": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_int, POINTER(idrs_krylov_opts_t),
                                   c_double, c_int, c_void_p, c_void_p, POINTER(idrs_stats_t), POINTER(c_int),
                                   POINTER(c_int)]),
    "idrs_msqmridr_c128": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, POINTER(idrs_krylov_opts_t),
                                   c_void_p, c_void_p, POINTER(idrs_stats_t)]),
    "idrs_idrs_mixed_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_opts_t), c_double, c_int, c_int,
                                    c_void_p, POINTER(idrs_stats_t), POINTER(c_int)]),
    "idrs_idrs_mixed_c128": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(idrs_opts_t), c_double, c_int, c_int,
                                     c_void_p, POINTER(idrs_stats_t), POINTER(c_int)]),
    "idrs_ms_scaleback_f64": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int]),
    "idrs_ms_scaleback_c128": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int]),
    "idrs_crs_transpose": (c_int, [c_void_p, POINTER(c_void_p)]),
    "idrs_crs_transpose_refresh": (c_int, [c_void_p, c_void_p]),
    "idrs_user_matrix_set_values_f64": (c_int, [c_void_p, c_void_p, c_int]),
    "idrs_crs_pattern_grad_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "idrs_batch_format": (c_int, [c_int, c_void_p, POINTER(c_void_p)]),
    "idrs_batch_info": (c_int, [c_void_p, POINTER(c_int), c_void_p]),
    "idrs_idrs_batched_mixed_f64": (c_int, [c_void_p, c_void_p, c_int, POINTER(idrs_opts_t), c_double, c_int, c_void_p,
                                            c_void_p, c_void_p]),
    "idrs_idrs_batched_f64": (c_int, [c_void_p, c_void_p, c_int, POINTER(idrs_opts_t), c_void_p, c_void_p]),
    "idrs_idrs_batched_c128": (c_int, [c_void_p, c_void_p, c_int, POINTER(idrs_opts_t), c_void_p, c_void_p]),
    "idrs_make_p_f64": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_uint64, c_void_p]),
    "idrs_make_p_c128": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_uint64, c_void_p]),
}

_lib = None


def _init_torch_runtime_first():
    """torch ships a HIP runtime of its own beside the system one that
    libidrs_hip.so links.  With both in one process, torch finds no device
    if the system runtime initialised the GPU first, so torch's runtime is
    brought up before the library is loaded (device tensors then pass
    between the two as plain device addresses)."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.device_count() > 0:
            torch.cuda.init()
    except RuntimeError:
        pass


def load():
    """Load libidrs_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    _init_torch_runtime_first()
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_PROTOS)


def check(status):
    if status != IDRS_OK:
        msg = load().idrs_last_error()
        raise IdrsError(status, msg.decode() if msg else "")
    return status
