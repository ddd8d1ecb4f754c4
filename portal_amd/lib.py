"""ctypes binding of the C ABI in ``include/portal_hip.h`` (``libportal_hip.so``).

The library is built in-tree by ``portal_amd.build`` (or ``__graft_entry__.build()``).
There is no CPU fallback: if the library is missing, or no GPU is visible when a
context is created, these calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libportal_hip.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

PORTAL_FP64 = 0
PORTAL_MIXED = 1

STATUS = {0: "PORTAL_OK", -1: "PORTAL_ERR_ARG", -2: "PORTAL_ERR_DEVICE", -3: "PORTAL_ERR_DEPTH",
          -4: "PORTAL_ERR_DEPTH_RT", -5: "PORTAL_ERR_LU", -6: "PORTAL_ERR_TEMP", -7: "PORTAL_ERR_STATE",
          -8: "PORTAL_ERR_NOMEM", -9: "PORTAL_ERR_UNSUPPORTED"}

# every entry point declared in include/portal_hip.h
EXPORTS = ["portal_ctx_create", "portal_ctx_destroy", "portal_last_error", "portal_upload_grid",
           "portal_upload_species", "portal_species_info", "portal_set_levpop", "portal_get_levpop",
           "portal_se_pass", "portal_inverse_rt", "portal_inverse_rt_device", "portal_se_solve",
           "portal_lambda_iterate",
           "portal_image_pass", "portal_mlp_create", "portal_mlp_destroy", "portal_mlp_num_params",
           "portal_mlp_set_params", "portal_mlp_get_params", "portal_mlp_forward", "portal_mlp_train_step",
           "portal_mlp_train", "portal_surrogate_features", "portal_get_pol_pops",
           "portal_batch_lambda_iterate", "portal_depol_table", "portal_set_depol"]


class SeParams(C.Structure):
    _fields_ = [("n_sphere", C.c_int), ("xth", dp), ("nvel", C.c_int), ("dx", C.c_double), ("gvec", dp),
                ("bvec_all", dp), ("n_depth", C.c_int), ("extra_info", C.c_int), ("precision", C.c_int),
                ("star_mod", C.c_int), ("nth", C.c_int), ("nph", C.c_int), ("R_star", C.c_double),
                ("T_star", C.c_double), ("X_star", dp)]


class IterParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol_pop", C.c_double), ("tol_align", C.c_double),
                ("surrogate", C.c_int), ("fallback_tol", C.c_double), ("rebuild_paths", C.c_int),
                ("ng_accel", C.c_int), ("ali", C.c_int)]


class ImgParams(C.Structure):
    _fields_ = [("nlevels", C.c_int), ("l1", ip), ("l2", ip), ("nu0_all", dp), ("Aji_all", dp), ("kd_all", dp),
                ("th_inc", C.c_double), ("ph_az", C.c_double), ("nvel", C.c_int), ("vel_out", dp),
                ("im_pix", C.c_int), ("im_size", C.c_double), ("dx", C.c_double), ("alpha", C.c_double),
                ("n_depth_rt", C.c_int), ("precision", C.c_int), ("star_mod", C.c_int), ("R_star", C.c_double),
                ("T_star", C.c_double), ("X_star", dp), ("node_pixels", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("t_walk_ms", C.c_double), ("t_prep_ms", C.c_double), ("t_rt_ms", C.c_double),
                ("t_se_ms", C.c_double), ("t_img_ms", C.c_double), ("t_total_ms", C.c_double),
                ("rt_kernel_avg_ms", C.c_double), ("n_cross", C.c_longlong), ("n_paths", C.c_longlong),
                ("iterations", C.c_int), ("fallbacks", C.c_int), ("lu_failures", C.c_int),
                ("max_rel_change", C.c_double), ("rt_variant", C.c_int), ("rt_waves", C.c_int),
                ("max_path", C.c_int), ("ng_steps", C.c_int), ("converged", C.c_int),
                ("max_align_change", C.c_double), ("se_form", C.c_int), ("se_waves", C.c_int),
                ("t_mlp_ms", C.c_double), ("rt_order", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m portal_amd.build` "
                           "(there is no CPU fallback for the PORTAL hot path)")
    lib = C.CDLL(path)
    lib.portal_last_error.restype = C.c_char_p
    lib.portal_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    for n in EXPORTS:
        f = getattr(lib, n)
        if n not in ("portal_last_error",):
            f.restype = C.c_int
    lib.portal_inverse_rt_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.portal_mlp_num_params.restype = C.c_longlong
    lib.portal_mlp_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]
    lib.portal_mlp_train_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p,
                                          C.c_void_p]
    lib.portal_mlp_train.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float,
                                     C.c_ulonglong, C.c_void_p, C.c_void_p]
    _lib = lib
    return lib
