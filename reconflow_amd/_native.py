"""ctypes bindings of the native libraries built by :mod:`reconflow_amd.build`.

Loading is strict: a missing ``librfhip.so`` raises instead of falling back to any other
implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

BUILD = Path(__file__).resolve().parent / "_build"


class RFParams(C.Structure):
    """Mirror of ``rf_params`` (include/rfhip.h)."""

    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("channels", C.c_int), ("input_type", C.c_int),
        ("n_levels", C.c_int), ("scale", C.c_float), ("method", C.c_int),
        ("lambda_", C.c_float), ("lambdagrad", C.c_float), ("theta", C.c_float), ("tau", C.c_float),
        ("n_warp_iters", C.c_int), ("n_solver_iters", C.c_int), ("with_visualization", C.c_int),
        ("lambdaf", C.c_float), ("lambdams", C.c_float), ("alpha_tv", C.c_float), ("alpha_proj", C.c_float),
        ("alpha_fn", C.c_float), ("stride_alignment", C.c_int), ("path", C.c_int), ("storage", C.c_int),
        ("use_graph", C.c_int), ("fix_coarse_k", C.c_int), ("tri_f64", C.c_int),
        ("joint_tile", C.c_int), ("flow_tile", C.c_int),
    ]


_P = C.c_void_p
_F = C.c_float
_I = C.c_int
_D = C.POINTER(C.c_double)

# name -> (restype, argtypes)
_PROTOS = {
    "rf_params_default": (None, [C.POINTER(RFParams), _I, _I]),
    "rf_last_error": (C.c_char_p, []),
    "rf_compute_pyramid_levels": (_I, [_I, _I, _I, _F]),
    "rf_create": (_I, [C.POINTER(_P), _I, _I, _I]),
    "rf_destroy": (_I, [_P]),
    "rf_initialize": (_I, [_P, C.POINTER(RFParams)]),
    "rf_initialize_large_displacement": (_I, [_P]),
    "rf_get_pyramid": (_I, [_P, _P, _P, _P]),
    "rf_get_stride": (_I, [_P]),
    "rf_get_stream": (_P, [_P]),
    "rf_copy_images": (_I, [_P, _P, _P]),
    "rf_copy_images_device": (_I, [_P, _P, _P]),
    "rf_set_camera": (_I, [_P, _D, _D]),
    "rf_copy_initial_flow": (_I, [_P, _P, _P]),
    "rf_copy_sparse_flow": (_I, [_P, _P, _P, _P]),
    "rf_copy_warm_start": (_I, [_P, _P, _P]),
    "rf_warm_start_in_use": (_I, [_P]),
    "rf_solve_flow": (_I, [_P, _F]),
    "rf_solve_flow_noalloc": (_I, [C.POINTER(RFParams), _P, _P, _P, _P, _P]),
    "rf_warmnet_create": (_I, [C.POINTER(_P), _I, _I, _I]),
    "rf_warmnet_destroy": (_I, [_P]),
    "rf_warmnet_param_count": (_I, [_P]),
    "rf_warmnet_set_params": (_I, [_P, _P, _I]),
    "rf_warmnet_get_params": (_I, [_P, _P, _I]),
    "rf_warmnet_get_grads": (_I, [_P, _P, _I]),
    "rf_warmnet_input": (_I, [_P, _P, _P, _I, C.c_long, _P, _P, _I, C.c_long, _I, _P]),
    "rf_warmnet_forward": (_I, [_P, _I, _P, _P]),
    "rf_warmnet_train_step": (_I, [_P, _I, _P, _F, C.POINTER(_F), _P]),
    "rf_warmnet_warm_start": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P]),
    "rf_copy_masks": (_I, [_P, _P, _P]),
    "rf_copy_depth": (_I, [_P, _P, _P]),
    "rf_solve_inpainting": (_I, [_P, _F]),
    "rf_solve_scene_flow": (_I, [_P, _F, _F, _F, _F, _F, _F, _I]),
    "rf_copy_scene_flow": (_I, [_P, _P, _P, _P, _P]),
    "rf_solve_recon": (_I, [_P, _D, _D, _D, _D, _F]),
    "rf_synchronize": (_I, [_P]),
    "rf_copy_flow": (_I, [_P, _P, _P, _P]),
    "rf_copy_3d": (_I, [_P, _P, _P, _P]),
    "rf_get_result_device": (_I, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "rf_remove_edge_artifacts_3d": (_I, [_P, _I, _F]),
    "rf_read_flo": (_I, [C.c_char_p, C.POINTER(_I), C.POINTER(_I), _P, _P, _I]),
    "rf_write_flo": (_I, [C.c_char_p, _P, _P, _I, _I, _I]),
    "rf_write_ply": (_I, [C.c_char_p, _P, _P, _P, _P, _I, _I, _I, _F, _F, _F]),
    "rf_profile_arm": (_I, [_P, _I, _I]),
    "rf_profile_read": (_I, [_P, C.POINTER(C.c_double), C.POINTER(_I)]),
    "rf_joint_tiling": (_I, [_P, C.POINTER(_I), _I]),
    "rf_flow_tiling": (_I, [_P, C.POINTER(_I), _I]),
    "rf_batch_create": (_I, [C.POINTER(_P), C.POINTER(RFParams), _I, _P, _P]),
    "rf_batch_destroy": (_I, [_P]),
    "rf_batch_stride": (_I, [_P, _I]),
    "rf_batch_set_images_device": (_I, [_P, _I, _P, _P]),
    "rf_batch_set_camera": (_I, [_P, _I, _D, _D, _D, _D]),
    "rf_batch_solve": (_I, [_P]),
    "rf_batch_synchronize": (_I, [_P]),
    "rf_batch_get_result_device": (_I, [_P, _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                        C.POINTER(_P)]),
    "rf_batch_get_stream": (_P, [_P]),
    "rf_batch_set_warm_start_device": (_I, [_P, _I, _P, _P]),
    "rf_batch_warm_start_in_use": (_I, [_P, _I]),
    "rf_batch_tiling": (_I, [_P, C.POINTER(_I), _I]),
    # per-kernel entry points (device pointers as ints, stream last)
    "rf_k_rgb_to_gray": (_I, [_P, _P, _I, _I, _I, _P]),
    "rf_k_cv8u_to_gray": (_I, [_P, _P, _I, _I, _I, _P]),
    "rf_k_cv16u_to_gray": (_I, [_P, _P, _I, _I, _I, _P]),
    "rf_k_cv32f_to_gray": (_I, [_P, _P, _I, _I, _I, _P]),
    "rf_k_downscale": (_I, [_P, _I, _I, _I, _I, _I, _I, _P, _P]),
    "rf_k_downscale_scaled": (_I, [_P, _I, _I, _I, _I, _I, _I, _F, _P, _P]),
    "rf_k_upscale": (_I, [_P, _I, _I, _I, _I, _I, _I, _F, _P, _P]),
    "rf_k_warp_image": (_I, [_P, _I, _I, _I, _P, _P, _P, _P]),
    "rf_k_compute_derivatives": (_I, [_P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "rf_k_solve_data_l1": (_I, [_P] * 9 + [_I, _I, _I, _F, _F, _P, _P, _P]),
    "rf_k_solve_smooth_dual_tv_global": (_I, [_P] * 6 + [_I, _I, _I, _F, _F] + [_P] * 5),
    "rf_k_median_filter": (_I, [_P, _P, _I, _I, _I, _P, _P, _I, _P]),
    "rf_k_median_filter_3d": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _I, _P]),
    "rf_k_add": (_I, [_P, _P, _I, _P, _P]),
    "rf_k_uniform_divide": (_I, [_P, _I, _F, _P, _P]),
    "rf_k_solve_recon_data_l1": (_I, [_P] * 15 + [_I, _I, _I, _F, _F, _F, _P, _P, _P]),
    "rf_k_solve_recon_data_l1_fn": (_I, [_P] * 17 + [_I, _I, _I, _F, _F, _F, _F, _P, _P, _P, _P]),
    "rf_k_solve_uvproj": (_I, [_P] * 8 + [_I, _I, _I, _F, _F, _P, _P, _P]),
    "rf_k_solve_3d_minimal_area": (_I, [_P] * 5 + [_D, _D, _F, _F, _I, _I, _I, _P, _P, _P, _P]),
    "rf_k_cleanup_3d": (_I, [_P] * 6 + [_I, _I, _I, _P, _P, _P, _P]),
    "rf_k_solve_uc": (_I, [_P, _P, _P, _D, _D, _D, _I, _I, _I, _P, _P, _P]),
    "rf_k_update_sk": (_I, [_P] * 8 + [_I, _I, _I, _P, _P, _P]),
    "rf_k_flow_to_hsv": (_I, [_P, _P, _I, _I, _I, _P, _F, _P]),
    "rf_k_remove_edge_artifacts_3d": (_I, [_P, _P, _P, _I, _F, _I, _I, _I, _P, _P, _P, _P]),
    "rf_k_compute_deriv": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "rf_k_compute_derivatives_patch": (_I, [_P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "rf_k_solve_data_charb_for_tv": (_I, [_P] * 17 + [_I, _I, _I, _F, _F, _F] + [_P] * 5),
    "rf_k_solve_data_l1_inpaint": (_I, [_P] * 11 + [_I, _I, _I, _F, _F, _P, _P, _P]),
    "rf_k_cv16u_to_32f": (_I, [_P, _P, _I, _I, _I, _P]),
    "rf_k_solve_scene_flow": (_I, [_P, _P, _P, _P, _F, _F, _F, _F, _I, _I, _I, _P, _P, _P, _P]),
    "rf_k_scene_flow_to_hsv": (_I, [_P, _I, _I, _I, _P, _F, _P]),
}

_SYNTH_PROTOS = {
    "rfs_make_flow_pair": (_I, [_I, _I, _I, C.c_uint64, _P, _P, _P, _P]),
    "rfs_make_recon_pair": (_I, [_I, _I, _I, C.c_uint64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "rfs_make_sequence": (_I, [_I, _I, _I, C.c_uint64, _I, _P, _P, _P, _P]),
}

_lib = None
_synth = None


def _bind(lib: C.CDLL, protos: dict) -> C.CDLL:
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib_path() -> Path:
    """The in-tree build, or the library named by RFHIP_LIBRARY (another build of the same ABI)."""
    alt = os.environ.get("RFHIP_LIBRARY")
    return Path(alt) if alt else BUILD / "librfhip.so"


def lib() -> C.CDLL:
    """The HIP solver library; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run reconflow_amd.build.build_all() (HIP extension not built)")
        # One HIP runtime per process: torch ships its own libamdhip64 (same soname). Loading it
        # first lets the solver bind to it; loading the solver first would bring a second
        # runtime, and device pointers of one are invalid in the other.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = _bind(C.CDLL(str(path)), _PROTOS)
    return _lib


def synth() -> C.CDLL:
    global _synth
    if _synth is None:
        path = BUILD / "librfsynth.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run reconflow_amd.build.build_all()")
        _synth = _bind(C.CDLL(str(path)), _SYNTH_PROTOS)
    return _synth


def declared_symbols() -> list[str]:
    return list(_PROTOS)


def last_error() -> str:
    return lib().rf_last_error().decode(errors="replace")


class RFError(RuntimeError):
    pass


def check(status: int, what: str) -> None:
    if status != 0:
        raise RFError(f"{what} failed with status {status}: {last_error()}")
