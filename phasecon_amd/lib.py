"""ctypes binding of the in-tree native library phasecon_amd/libmbir4d.so.

The library exports the drop-in ``reconstruct`` (include/mbir4d.h), the
engine helpers and ``reconstruct_ex`` (include/mbir_engine.h) and the HIP
C-ABI (include/mbir_hip.h).  There is no fallback: if the library is not
built, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmbir4d.so")

K_AMATRIX, K_FWDPROJ, K_SWEEP, K_COST, K_BACKPROJ = range(5)


class HipGeom(C.Structure):
    """mirror of mbir_hip_geom (include/mbir_hip.h)"""
    _fields_ = [(n, C.c_int32) for n in ("N_r", "N_t", "N_p", "N_x", "N_y", "N_z",
                                          "N_time", "mult_xy", "mult_z")] + \
        [(n, C.c_double) for n in ("delta_r", "delta_t", "R0", "RMax", "x0", "y0",
                                   "delta_xy", "OffsetR", "q", "p_s", "p_t", "c_s",
                                   "c_t", "sig_s_q", "sig_s_qp", "sig_t_q", "sig_t_qp")] + \
        [("filt", C.c_double * 27), ("filt_t", C.c_double), ("alpha", C.c_double),
         ("huber_T", C.c_double), ("huber_delta", C.c_double), ("var_est", C.c_double)]


class GuardState(C.Structure):
    """mirror of mbir_guard (include/mbir_engine.h)"""
    _fields_ = [("best", C.c_double), ("tol", C.c_double), ("recoveries", C.c_int32),
                ("max_recoveries", C.c_int32)]


GUARD_OK, GUARD_SAVE, GUARD_CONFIRM, GUARD_RECOVER, GUARD_FAIL = range(5)


class PcPrior(C.Structure):
    """mirror of mbir_hip_pc_prior (include/mbir_hip.h)"""
    _fields_ = [("mu", C.c_double), ("wavenumber", C.c_double), ("alpha", C.c_double),
                ("T", C.c_double * 4)] + \
        [(n, C.c_double * 2) for n in ("c_s", "sig_s_q", "sig_s_qp", "c_t", "sig_t_q",
                                       "sig_t_qp")] + \
        [("p", C.c_double), ("q", C.c_double), ("filt", C.c_double * 27), ("filt_t", C.c_double)]


class Options(C.Structure):
    """mirror of mbir_options (include/mbir_engine.h)"""
    _fields_ = [("p_s", C.c_double), ("p_t", C.c_double), ("tile", C.c_int32),
                ("alpha", C.c_double), ("max_equits", C.c_int32),
                ("single_stage", C.c_int32), ("write_files", C.c_int32),
                ("stop_cost", C.c_double), ("seed", C.c_uint32), ("refresh_every", C.c_int32)]


_LIB = None

# name -> (restype, argtypes); None argtypes = unchecked pointers
_SIGS = {
    "reconstruct": (C.c_int, None),
    "reconstruct_ex": (C.c_int, None),
    "mbir_default_options": (None, [C.POINTER(Options)]),
    "mbir_multires_schedule": (C.c_int32, None),
    "mbir_view_map": (C.c_int32, None),
    "mbir_stage_geometry": (C.c_int, None),
    "mbir_detector_response": (None, None),
    "mbir_sincos": (None, None),
    "mbir_update_mask": (None, None),
    "mbir_sino_to_internal": (None, None),
    "mbir_hip_last_error": (C.c_char_p, []),
    "mbir_hip_create": (C.c_int, None),
    "mbir_hip_destroy": (None, [C.c_void_p]),
    "mbir_hip_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "mbir_hip_set_object": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_get_object": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_upsample_object": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "mbir_hip_build_amatrix": (C.c_int, [C.c_void_p]),
    "mbir_hip_get_amatrix": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mbir_hip_init_error": (C.c_int, [C.c_void_p]),
    "mbir_hip_get_error": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_set_error": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_forward_project": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mbir_hip_backproject": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mbir_hip_icd_equits": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_uint32, C.c_void_p]),
    "mbir_hip_icd_graph": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_uint32]),
    "mbir_hip_nhicd": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_uint32,
                                 C.c_void_p]),
    "mbir_hip_get_magnitudes": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mbir_hip_phantom_project": (C.c_int, [C.c_void_p] * 2 + [C.c_int32] * 3 + [C.c_void_p]
                                 + [C.c_int32] * 2 + [C.c_double] * 3 + [C.c_void_p] * 2),
    "mbir_hip_std_downsample": (C.c_int, [C.c_void_p] * 2 + [C.c_int32] * 4
                                + [C.c_double, C.c_void_p, C.c_void_p]),
    "mbir_hip_aps_normalize": (C.c_int, [C.c_void_p] * 3 + [C.c_int32] * 6
                               + [C.c_double, C.c_void_p, C.c_void_p]),
    "mbir_interlaced_views": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_int32, C.c_float, C.c_float]),
    "mbir_hip_offset_update": (C.c_int, [C.c_void_p, C.c_int32]),
    "mbir_hip_get_offset": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_variance_update": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_get_weight": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_cost": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "mbir_hip_substitution_batch": (C.c_int, None),
    "mbir_hip_substitution_batch_noalloc": (C.c_int, None),
    "mbir_hip_voxel_thetas": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mbir_hip_kernel_time": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "mbir_hip_reset_kernel_time": (C.c_int, [C.c_void_p]),
    "mbir_hip_xcd_check": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mbir_hip_object_snapshot": (C.c_int, [C.c_void_p, C.c_int32]),
    "mbir_hip_pc_enable": (C.c_int, [C.c_void_p, C.POINTER(PcPrior)]),
    "mbir_hip_pc_set_objects": (C.c_int, [C.c_void_p] * 3),
    "mbir_hip_pc_get_objects": (C.c_int, [C.c_void_p] * 3),
    "mbir_hip_pc_init_error": (C.c_int, [C.c_void_p] * 3),
    "mbir_hip_pc_get_error": (C.c_int, [C.c_void_p] * 3),
    "mbir_hip_pc_equits": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_uint32, C.c_void_p]),
    "mbir_hip_pc_cost": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_pc_admm_init": (C.c_int, [C.c_void_p] * 3 + [C.c_int32]),
    "mbir_hip_pc_admm_target": (C.c_int, [C.c_void_p] * 3),
    "mbir_hip_pc_admm_update": (C.c_int, [C.c_void_p] * 4),
    "mbir_hip_pc_admm_state": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "mbir_guard_init": (None, [C.POINTER(GuardState), C.c_double, C.c_double, C.c_int32]),
    "mbir_guard_check": (C.c_int32, [C.POINTER(GuardState), C.c_double, C.c_int32]),
    "mbir_guard_alpha": (C.c_double, [C.c_double]),
    "mbir_hip_synchronize": (C.c_int, [C.c_void_p]),
    "mbir_hip_bp_normalized": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mbir_hip_fbp": (C.c_int, [C.c_void_p, C.c_void_p]),
    # learned warm start (include/unet_hip.h)
    "unet_last_error": (C.c_char_p, []),
    "unet_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_uint32]),
    "unet_destroy": (None, [C.c_void_p]),
    "unet_param_count": (C.c_int64, [C.c_void_p]),
    "unet_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "unet_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "unet_forward": (C.c_int, [C.c_void_p] + [C.c_int32] * 4 + [C.c_void_p, C.c_void_p]),
    "unet_gradients": (C.c_int, [C.c_void_p] + [C.c_int32] * 4 + [C.c_void_p] * 4),
    "unet_train_step": (C.c_int, [C.c_void_p] + [C.c_int32] * 4 +
                        [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]),
    "unet_conv3d": (C.c_int, [C.c_int32] * 7 + [C.c_void_p] * 3 + [C.c_int32] +
                    [C.c_void_p] * 2),
    "unet_mfma_probe": (C.c_int, [C.c_void_p] * 3),
}


def exported_names():
    return list(_SIGS)


def load():
    """Loads libmbir4d.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m phasecon_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            if args is not None:
                fn.argtypes = args
        _LIB = L
    return _LIB


def last_error():
    return load().mbir_hip_last_error().decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {last_error()}")
