"""ctypes binding of libmaxlq.so (include/maxlq.h).

The product path: every call goes to the HIP library; there is no CPU
fallback.  ``load()`` raises if the in-tree library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmaxlq.so")

MLQ_OK, MLQ_ENAN, MLQ_EBADKIND, MLQ_EHIP, MLQ_EOOM, MLQ_EINVAL, MLQ_EUNSUP = range(7)
MLQ_FP64, MLQ_MIXED = 0, 1
MLQ_IP_L2, MLQ_IP_HL, MLQ_IP_H, MLQ_IP_OLD_H, MLQ_IP_OLD_HL = range(5)
MLQ_J_MAXDLQDT, MLQ_J_LINEMIN, MLQ_J_LQNORM = range(3)
MLQ_LOG_ITER, MLQ_LOG_LINEMIN, MLQ_LOG_BRENT, MLQ_LOG_MSG, MLQ_LOG_ERR = 1, 2, 3, 4, 5
MLQ_LOG_SPECTRUM = 6
MLQ_LOG_FIELD = 7
MLQ_LOG_KAPPA = 8
MLQ_RCG_KAPPA = 4
MLQ_PERT_SINE, MLQ_PERT_DIVFREE_SINE, MLQ_PERT_LOAD_RAW, MLQ_PERT_LOAD = 0, 1, 2, 3
MLQ_PERT_LOAD_SMOOTH, MLQ_PERT_LOAD_SMOOTH_RAW, MLQ_PERT_LOAD_K = 4, 5, 6
MLQ_PERT_LOAD_POLY, MLQ_PERT_LOAD_EXP, MLQ_PERT_RANDOM_POLY, MLQ_PERT_RANDOM = 7, 8, 9, 10
MLQ_RCG_REFERENCE = 1   # mlq_rcg_cfg.flags: reference-structured iteration
MLQ_RCG_LEAN = 8        # mlq_rcg_cfg.flags: line search from coefficient fields

# symbols the header declares (checked by the CPU test suite)
EXPORTS = [
    "mlq_default_params", "mlq_init", "mlq_finalize", "mlq_set_params", "mlq_last_error",
    "mlq_local_size_3d", "mlq_sync", "mlq_fftfwd", "mlq_fftbwd", "mlq_fftfwdr", "mlq_fftbwdr",
    "mlq_fftfwdV", "mlq_fftbwdV", "mlq_fftfwdr_noalloc", "mlq_gradient", "mlq_divergence", "mlq_laplacian",
    "mlq_solve_poisson", "mlq_dealias_scalar", "mlq_divAvg_free", "mlq_HilbertGradient",
    "mlq_calc0thFourierModes", "mlq_calc_uk", "mlq_calc_gk_order2", "mlq_calc_global_Lq_norm",
    "mlq_calculate_spectral_data", "mlq_spectral_data_field", "mlq_calc_global_summed_enstrophy",
    "mlq_calc_global_dEnstrophydt",
    "mlq_rescaleLqNorm", "mlq_calcConstraintDerivativeL2", "mlq_vectorTransport",
    "mlq_inner_product", "mlq_global_inner_product", "mlq_global_summed_field_inner_product",
    "mlq_calc_local_dLqdt_inclParts", "mlq_calc_global_dLqdt_inclParts", "mlq_GradL2ForLq",
    "mlq_eval_J", "mlq_field_alloc", "mlq_field_free", "mlq_upload", "mlq_download",
    "mlq_seeded_field", "mlq_eval_J_batch", "mlq_time_linesearch", "mlq_kappa_test", "mlq_rcg_path", "mlq_rcg_solve", "mlq_time_fft", "mlq_time_copy", "mlq_time_column_copy",
    "mlq_autotune", "mlq_set_tiles", "mlq_sizes",
    "mlq_prof_enable", "mlq_prof_get", "mlq_prof_name", "mlq_mem_info", "mlq_mem_live",
    "mlq_transform_counts", "mlq_save_field_nc", "mlq_read_field_nc", "mlq_result_filename",
    "mlq_q_text", "mlq_parse_result_filename", "mlq_fftw_mpi_local_size_3d",
    "mlq_diagnosticScalars", "mlq_kappa_test_pert",
]


class MlqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"maxlq error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("q", C.c_double), ("B", C.c_double), ("visc", C.c_double),
                ("lambda1", C.c_double), ("viscCoefficient", C.c_double),
                ("pressureCoefficient", C.c_double), ("toDealias", C.c_int),
                ("use_e_u", C.c_int), ("dealiase_if_mult_by_e_u", C.c_int),
                ("precision", C.c_int)]


class RcgCfg(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("optim_tol", C.c_double), ("start_iter", C.c_int),
                ("batch_linesearch", C.c_int), ("reset_every", C.c_int), ("rolling", C.c_int),
                ("flags", C.c_int), ("spectra_every", C.c_int), ("save_every", C.c_int),
                ("kappa_every", C.c_int), ("kappa_seed", C.c_ulonglong),
                ("kappa_field", C.POINTER(C.c_double))]


class RcgResult(C.Structure):
    _fields_ = [("J1", C.c_double), ("iters", C.c_int), ("successful", C.c_int),
                ("while_flag", C.c_int), ("stop_reason", C.c_int), ("n_J", C.c_int),
                ("n_J_computed", C.c_int), ("seconds", C.c_double), ("n_J_poly", C.c_int),
                ("pad_", C.c_int), ("enstrophy", C.c_double), ("dEnstrophydt", C.c_double)]


LOG_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                     C.POINTER(C.c_double))

_lib = None


def load(path: str = LIB_PATH):
    """Load the in-tree HIP library; raise loudly when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built: run `python -m maxlq_amd.build` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    ci, cd = C.c_int, C.c_double
    sig = {
        "mlq_default_params": (None, [C.POINTER(Params), cd]),
        "mlq_init": (ci, [C.POINTER(vp), ci, ci, ci, C.POINTER(Params)]),
        "mlq_finalize": (ci, [vp]),
        "mlq_set_params": (ci, [vp, C.POINTER(Params)]),
        "mlq_last_error": (C.c_char_p, [vp]),
        "mlq_local_size_3d": (ci, [vp, C.POINTER(ci), C.POINTER(ci)]),
        "mlq_sync": (ci, [vp]),
        "mlq_fftfwd": (ci, [vp, dp, dp]), "mlq_fftbwd": (ci, [vp, dp, dp]),
        "mlq_fftfwdr": (ci, [vp, dp, dp]), "mlq_fftbwdr": (ci, [vp, dp, dp]),
        "mlq_fftfwdV": (ci, [vp, dp, dp]), "mlq_fftbwdV": (ci, [vp, dp, dp]),
        "mlq_fftfwdr_noalloc": (ci, [vp, vp, vp]),
        "mlq_gradient": (ci, [vp, dp, dp]), "mlq_divergence": (ci, [vp, dp, dp]),
        "mlq_laplacian": (ci, [vp, dp]), "mlq_solve_poisson": (ci, [vp, dp, cd, dp]),
        "mlq_dealias_scalar": (ci, [vp, dp, cd]), "mlq_divAvg_free": (ci, [vp, dp]),
        "mlq_HilbertGradient": (ci, [vp, dp, cd]), "mlq_calc0thFourierModes": (ci, [vp, dp, dp]),
        "mlq_calc_uk": (ci, [vp, dp, cd, dp]), "mlq_calc_gk_order2": (ci, [vp, dp, cd, dp]),
        "mlq_calc_global_Lq_norm": (ci, [vp, dp, dp]),
        "mlq_calculate_spectral_data": (ci, [vp, dp, ci, dp]),
        "mlq_spectral_data_field": (ci, [vp, ci, ci, dp]),
        "mlq_calc_global_summed_enstrophy": (ci, [vp, dp, dp]),
        "mlq_calc_global_dEnstrophydt": (ci, [vp, dp, dp]), "mlq_rescaleLqNorm": (ci, [vp, dp, cd]),
        "mlq_calcConstraintDerivativeL2": (ci, [vp, dp, dp]),
        "mlq_vectorTransport": (ci, [vp, dp, cd, cd, dp, dp, dp]),
        "mlq_inner_product": (ci, [vp, dp, dp, ci, dp]),
        "mlq_global_inner_product": (ci, [vp, dp, dp, ci, dp]),
        "mlq_global_summed_field_inner_product": (ci, [vp, dp, dp, ci, dp]),
        "mlq_calc_local_dLqdt_inclParts": (ci, [vp, dp, cd, dp]),
        "mlq_calc_global_dLqdt_inclParts": (ci, [vp, dp, cd, dp]),
        "mlq_GradL2ForLq": (ci, [vp, dp, dp]),
        "mlq_eval_J": (ci, [vp, dp, ci, dp]),
        "mlq_field_alloc": (ci, [vp, ci, C.POINTER(ci)]), "mlq_field_free": (ci, [vp, ci]),
        "mlq_upload": (ci, [vp, ci, dp]), "mlq_download": (ci, [vp, ci, dp]),
        "mlq_seeded_field": (ci, [vp, ci, C.c_ulonglong, ci, cd]),
        "mlq_eval_J_batch": (ci, [vp, ci, ci, dp, ci, dp]),
        "mlq_time_linesearch": (ci, [vp, ci, ci, dp, ci, ci, dp, dp]),
        "mlq_rcg_path": (ci, [vp, C.POINTER(ci)]),
        "mlq_kappa_test": (ci, [vp, ci, ci, ci, ci, ci, dp, dp, dp, C.POINTER(ci)]),
        "mlq_rcg_solve": (ci, [vp, ci, C.POINTER(RcgCfg), C.POINTER(RcgResult), LOG_CB, vp]),
        "mlq_time_fft": (ci, [vp, ci, ci, dp]),
        "mlq_time_copy": (ci, [vp, ci, ci, dp]),
        "mlq_time_column_copy": (ci, [vp, ci, ci, ci, ci, dp, dp]),
        "mlq_autotune": (ci, [vp, ci, dp, C.POINTER(ci)]),
        "mlq_set_tiles": (ci, [vp, C.POINTER(ci)]),
        "mlq_sizes": (ci, [vp, dp, dp]),
        "mlq_prof_enable": (ci, [vp, ci]),
        "mlq_prof_get": (ci, [vp, ci, C.POINTER(C.c_longlong), dp, dp]),
        "mlq_prof_name": (C.c_char_p, [ci]),
        "mlq_mem_info": (ci, [vp, dp, dp, dp, dp]),
        "mlq_mem_live": (ci, [vp, dp, dp]),
        "mlq_save_field_nc": (ci, [C.c_char_p, ci, ci, ci, dp]),
        "mlq_diagnosticScalars": (ci, [vp, dp, dp, dp, dp, C.c_longlong,
                                       C.POINTER(C.c_longlong)]),
        "mlq_kappa_test_pert": (ci, [vp, ci, ci, cd, cd, cd, C.c_ulonglong, dp]),
        "mlq_read_field_nc": (ci, [C.c_char_p, ci, ci, ci, dp]),
        "mlq_result_filename": (ci, [C.c_char_p, ci, C.c_char_p, C.c_char_p, cd, ci, ci, ci]),
        "mlq_q_text": (ci, [cd, C.c_char_p, ci]),
        "mlq_parse_result_filename": (ci, [C.c_char_p, dp, C.POINTER(ci), C.POINTER(ci),
                                           C.POINTER(ci)]),
        "mlq_transform_counts": (ci, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(q=4.0, **kw) -> Params:
    lib = load()
    p = Params()
    lib.mlq_default_params(C.byref(p), float(q))
    for k, v in kw.items():
        setattr(p, k, v)
    return p
