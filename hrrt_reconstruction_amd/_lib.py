"""ctypes binding of libhrrt_osem.so (the C ABI declared in include/hrrt_osem.h).

The library is built in-tree by ``hrrt_reconstruction_amd.build`` (hipcc,
gfx950).  Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhrrt_osem.so")

HRRT_OK, HRRT_ERR_ARG, HRRT_ERR_HIP, HRRT_ERR_NOMEM, HRRT_ERR_UNSUPPORTED = range(5)


class HrrtError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """hrrt_params (scanner model + reconstruction options)."""
    _fields_ = [
        ("rings", C.c_int), ("crystal_radius", C.c_float), ("plane_sep", C.c_float),
        ("bin_size", C.c_float), ("def_elements", C.c_int), ("def_angles", C.c_int),
        ("x_pixels", C.c_int), ("span", C.c_int), ("maxdel", C.c_int), ("subsets", C.c_int),
        ("rel_fov", C.c_float), ("npskip", C.c_int), ("z_min", C.c_int), ("z_max", C.c_int),
        ("epsilon1", C.c_float), ("epsilon2", C.c_float), ("epsilon3", C.c_float),
        ("image_max", C.c_float),
    ]


class Dims(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("views", C.c_int), ("z_pixels", C.c_int), ("z4", C.c_int),
        ("segmentsize", C.c_int), ("th_pixels", C.c_int), ("sviews", C.c_int), ("subsets", C.c_int),
        ("spp", C.c_int), ("x_rebin", C.c_int), ("v_rebin", C.c_int),
        ("image_floats", C.c_longlong), ("sino_bins", C.c_longlong), ("file_bins", C.c_longlong),
        ("paircol_elems", C.c_longlong), ("fov_voxels", C.c_longlong), ("fov_bins", C.c_int),
    ]


# every entry point of include/hrrt_osem.h with its ctypes signature
_VP, _I = C.c_void_p, C.c_int


class ImageInfo(C.Structure):
    """hrrt_image_info: the contents of the image's Interfile header."""
    _fields_ = [("recontype", C.c_int), ("iterations", C.c_int)] + [
        (n, C.c_char_p) for n in ("true_file", "prompt_file", "delayed_file", "norm_file", "atten_file",
                                  "scatter_file", "sw_version", "sw_build_id")]


class DelaysParams(C.Structure):
    """hrrt_delays_params (gen_delays options)."""
    _fields_ = [("span", C.c_int), ("maxrd", C.c_int), ("lr_type", C.c_int), ("pitch", C.c_float),
                ("diam", C.c_float), ("thick", C.c_float), ("tau", C.c_float), ("lut_file", C.c_char_p)]


LM_NBLOCKS = 936


class LmOptions(C.Structure):
    """hrrt_lm_options (list-mode histogramming)."""
    _fields_ = [("mode", C.c_int), ("start_ms", C.c_int), ("end_ms", C.c_int), ("ignore_border", C.c_int),
                ("no_doi", C.c_int)]


class LmStats(C.Structure):
    """hrrt_lm_stats."""
    _fields_ = [(n, C.c_longlong) for n in ("prompts", "randoms", "events", "tags", "time_tags", "rejected",
                                            "transmission", "first_word", "end_word")] + [
        ("last_time_ms", C.c_int), ("block_total", C.c_longlong * LM_NBLOCKS),
        ("block_nsamples", C.c_int * LM_NBLOCKS), ("block_last", C.c_int * LM_NBLOCKS)]


SIGNATURES = {
    "hrrt_create": (_I, [C.POINTER(Params), C.POINTER(_VP)]),
    "hrrt_destroy": (None, [_VP]),
    "hrrt_get_dims": (_I, [_VP, C.POINTER(Dims)]),
    "hrrt_get_order": (_I, [_VP, _VP, _VP]),
    "hrrt_get_segments": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "hrrt_query_geometry": (_I, [C.POINTER(Params), C.POINTER(Dims)] + [_VP] * 6),
    "hrrt_set_norm_zero_mask": (_I, [_VP, _VP]),
    "hrrt_get_norm_zero_mask": (_I, [_VP, _VP]),
    "hrrt_last_error": (C.c_char_p, []),
    "hrrt_stage_sinograms": (_I, [_VP] * 9 + [_VP]),
    "hrrt_calculate_normfac": (_I, [_VP, _VP, _VP, _VP]),
    "hrrt_init_image": (_I, [_VP, _VP, _VP]),
    "hrrt_osem_subset": (_I, [_VP, _VP, _VP, _VP, _VP, _I, _VP]),
    "hrrt_forward_proj3d_noalloc": (_I, [_VP, _VP, _VP, _I, _VP]),
    "hrrt_stage_sinograms_w012": (_I, [_VP, _I, _I] + [_VP] * 8 + [_VP]),
    "hrrt_stage_sinograms_host": (_I, [_VP] * 9 + [_VP]),
    "hrrt_stage_sinograms_w012_host": (_I, [_VP, _I, _I] + [_VP] * 8 + [_VP]),
    "hrrt_osem_subset_w012": (_I, [_VP, _VP, _VP, _VP, _I, _VP]),
    "hrrt_end_iteration": (_I, [_VP, _VP, _VP]),
    "hrrt_osem_subset_frames": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _I, _VP]),
    "hrrt_osem_iteration_frames": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP]),
    "hrrt_to_half": (_I, [_VP, _VP, C.c_longlong, _VP]),
    "hrrt_to_half_range": (_I, [_VP, _VP, C.c_longlong, _VP, _VP]),
    "hrrt_calculate_normfac_mixed": (_I, [_VP, _VP, _VP, _VP]),
    "hrrt_osem_subset_mixed": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
    "hrrt_osem_subset_frames_mixed": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
    "hrrt_osem_iteration_frames_mixed": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "hrrt_count_moments": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "hrrt_timing_enable": (_I, [_VP, _I]),
    "hrrt_set_launch_shape": (_I, [_VP, _I, _I]),
    "hrrt_set_backproject_shape": (_I, [_VP, _I]),
    "hrrt_timing_read": (_I, [_VP, _VP, _VP]),
    "hrrt_forward_proj3d": (_I, [_VP, _VP, _VP, _I, _VP]),
    "hrrt_back_proj3d": (_I, [_VP, _VP, _VP, _I, _VP]),
    "hrrt_back_proj3d_incremental": (_I, [_VP, _VP, _VP, _I, _VP]),
    "hrrt_back_proj3d_rotframe": (_I, [_VP, _VP, _VP, _I, _I, _VP]),
    "hrrt_rotate_add": (_I, [_VP, _VP, _VP, _I, _VP]),
    "hrrt_set_backproject_schedule": (_I, [_VP, _I]),
    "hrrt_get_backproject_schedule": (_I, [_VP, _VP]),
    "hrrt_set_psf": (_I, [_VP, _I, C.c_float, C.c_float, C.c_float]),
    "hrrt_image_to_flat": (_I, [_VP, _VP, _VP, _I, _VP]),
    "hrrt_image_from_flat": (_I, [_VP, _VP, _VP, _I, _VP]),
    "hrrt_write_image": (_I, [_VP, _VP, C.c_char_p, _I, _VP, _VP]),
    "hrrt_read_image": (_I, [_VP, C.c_char_p, _VP, _I, _VP]),
    "hrrt_write_normfac": (_I, [_VP, _VP, C.c_char_p, _VP]),
    "hrrt_read_normfac": (_I, [_VP, C.c_char_p, _VP, _VP]),
    "hrrt_get_psf_kernel": (_I, [_VP, _I, _VP, _VP]),
    "hrrt_convolve3d": (_I, [_VP, _VP, _VP, _VP]),
    "hrrt_update_estimate_w3": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "hrrt_update_estimate_w012": (_I, [_VP, _VP, _VP, _VP]),
    "hrrt_cal_updateimage": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "hrrt_normfac_calculate": (_I, [_VP, _VP, _VP, _VP]),
    "hrrt_forward_proj3d_host": (_I, [_VP, _VP, _VP, _I]),
    "hrrt_back_proj3d_host": (_I, [_VP, _VP, _VP, _I]),
    "hrrt_back_proj3d_rotframe_host": (_I, [_VP, _VP, _VP, _I, _I]),
    "hrrt_rotate_add_host": (_I, [_VP, _VP, _VP, _I]),
    # include/hrrt_randoms.h
    "hrrt_delays_create": (_I, [C.POINTER(DelaysParams), C.POINTER(_VP)]),
    "hrrt_delays_destroy": (None, [_VP]),
    "hrrt_delays_get_dims": (_I, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "hrrt_delays_get_lut": (_I, [_VP, _I, _VP, _VP, _VP]),
    "hrrt_delays_save_lut": (_I, [_VP, C.c_char_p]),
    "hrrt_delays_singles": (_I, [_VP, _VP, C.c_float, _VP, _VP]),
    "hrrt_delays_smooth": (_I, [_VP, _VP, C.c_float, _VP, _VP]),
    "hrrt_gen_delays": (_I, [_VP, _VP, C.c_float, _VP, _VP, _VP, _VP]),
    "hrrt_delays_write": (_I, [_VP, _VP, C.c_char_p, _VP]),
    # include/hrrt_lmhist.h
    "hrrt_lm_histogram": (_I, [_VP, _VP, C.c_longlong, C.POINTER(LmOptions), _VP, _VP, _VP, _VP,
                               C.POINTER(LmStats), _VP]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libhrrt_osem.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run hrrt_reconstruction_amd.build.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(fn: str, code: int) -> None:
    if code != HRRT_OK:
        msg = lib().hrrt_last_error()
        raise HrrtError(fn, code, msg.decode() if msg else "")


def call(name: str, *args) -> int:
    code = getattr(lib(), name)(*args)
    check(name, code)
    return code
