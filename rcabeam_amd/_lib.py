"""ctypes binding of the MI355X shared library (the C ABI in include/rcabeam.h and
include/rcabeam_mi355x.h).

The library is loaded lazily on first use and there is no fallback: a missing
library or a missing GPU raises `RuntimeError` naming the cause.  Status codes
map to exceptions the way the reference bindings do
(src/rcabeam/bindings/python.cpp:15-25): argument errors become `ValueError`,
runtime errors `RuntimeError`.
"""

from __future__ import annotations

import ctypes
import os
import threading
from ctypes import c_char_p, c_double, c_float, c_int, c_size_t, c_ulonglong, c_void_p
from pathlib import Path

# RCABEAM_LIBRARY may name another build of the same ABI (e.g. a tuning variant).
LIBRARY_PATH = Path(os.environ.get("RCABEAM_LIBRARY", Path(__file__).resolve().parent / "lib" / "librcabeam_mi355x.so"))

SUCCESS, ERROR_DEVICE, ERROR_ARGUMENT = 0, 1, 2
STORAGE = {"float32": 0, "complex64": 0, "float16": 1, "bfloat16": 2, "float16-split": 3}
# rcabeam_coefficients (include/rcabeam_mi355x.h)
COEF_FLOAT32 = 0
COEF_16BIT = 1

# name -> argument types (all return rcabeam_status as int unless listed in _RESTYPE)
_SIGNATURES: dict[str, list] = {
    "rcabeam_status_string": [c_int],
    "rcabeam_opw_device": [c_void_p, c_void_p, c_size_t, c_size_t, c_size_t],
    "rcabeam_xdoppler_pd_device": [c_void_p, c_void_p] + [c_size_t] * 7,
    "rcabeam_rc_fmas_pd_device": [c_void_p, c_void_p] + [c_size_t] * 7,
    "rcabeam_delay_rca_channels_device": [c_void_p] * 6 + [c_size_t] * 4 + [c_int] + [c_float] * 4,
    "rcabeam_delay_rca_channel_data_device": [c_void_p] * 8 + [c_size_t] * 4 + [c_int] + [c_float] * 4,
    "rcabeam_opw_pd_from_channels_device": [c_void_p] * 8 + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_xdoppler_pd_from_channels_device": [c_void_p] * 8 + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_rc_fmas_pd_from_channels_device": [c_void_p] * 8 + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_fast_pd_from_channels_device": [c_void_p] * 10 + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_dmas_ccf_acf_from_channels_device": [c_void_p] * 8 + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_pack_ensemble_device": [c_void_p, c_void_p] + [c_size_t] * 4,
    "rcabeam_pack_ensemble_fp16_device": [c_void_p, c_void_p, c_void_p] + [c_size_t] * 4,
    "rcabeam_pack_ensemble_bf16_device": [c_void_p, c_void_p, c_void_p] + [c_size_t] * 4,
    "rcabeam_ensemble_device": [c_void_p] * 12 + [c_size_t] * 5 + [c_float] * 4 + [c_int, c_int],
    "rcabeam_opw_fp16_device": [c_void_p] * 10 + [c_size_t] * 5 + [c_float] * 4,
    "rcabeam_opw_bf16_device": [c_void_p] * 9 + [c_size_t] * 5 + [c_float] * 4,
    "rcabeam_opw_mixed_device": [c_void_p] * 9 + [c_int, c_int, c_int] + [c_size_t] * 5 + [c_float] * 4,
    "rcabeam_xdoppler_from_apertures_device": [c_void_p, c_size_t, c_size_t, c_void_p, c_void_p],
    "rcabeam_upload_host": [c_void_p, c_void_p, c_size_t, c_int],
    "rcabeam_upload_host_concurrent": [c_void_p, c_void_p, c_size_t, c_int],
    "rcabeam_opw_grid_aperture_device": [c_void_p, c_void_p] + [c_void_p] * 3 + [c_size_t] * 3 + [c_void_p] * 6
    + [c_int, c_int] + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_opw_grid_device": [c_void_p, c_void_p, c_int, c_int] + [c_void_p] * 3 + [c_size_t] * 3 + [c_void_p] * 6
    + [c_int, c_int] + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_unpack_ensemble_device": [c_void_p, c_void_p] + [c_size_t] * 4,
    "rcabeam_pack_ensemble_split_device": [c_void_p, c_void_p, c_void_p] + [c_size_t] * 4,
    "rcabeam_opw_grid_split_device": [c_void_p] * 6 + [c_size_t] * 3 + [c_void_p] * 6 + [c_int] + [c_size_t] * 4
    + [c_float] * 4,
    "rcabeam_opw_adjoint_grid_device": [c_void_p, c_void_p] + [c_void_p] * 3 + [c_size_t] * 3 + [c_void_p] * 8
    + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_opw_adjoint_workspace_bytes": [c_size_t] * 7,
    "rcabeam_opw_adjoint_grid_deterministic_device": [c_void_p, c_void_p] + [c_void_p] * 3 + [c_size_t] * 3
    + [c_void_p] * 8 + [c_void_p, c_size_t] + [c_size_t] * 4 + [c_float] * 4,
    "rcabeam_pack_ensemble_lowrank_device": [c_void_p, c_void_p, c_int, c_void_p, c_size_t, c_void_p] + [c_size_t] * 4,
    "rcabeam_lowrank_expand_device": [c_void_p, c_void_p, c_size_t, c_void_p, c_int, c_void_p, c_size_t, c_size_t],
    "rcabeam_dominant_subspace_device": [c_void_p, c_size_t, c_size_t, c_size_t, c_void_p],
    "rcabeam_count_active_taps_device": [c_void_p] * 5 + [c_size_t] * 4 + [c_float] * 4 + [c_void_p],
    "rcabeam_iq_absmax_device": [c_void_p, c_int, c_size_t, c_void_p, c_int],
    "rcabeam_gram_workspace_bytes": [c_size_t, c_size_t],
    "rcabeam_gram_real_elements": [c_size_t],
    "rcabeam_temporal_gram_device": [c_void_p, c_int, c_size_t, c_size_t, c_void_p, c_void_p, c_void_p, c_size_t, c_int],
    "rcabeam_temporal_gram_sampled_device": [c_void_p, c_int, c_size_t, c_size_t, c_size_t, c_void_p, c_void_p,
                                             c_void_p, c_size_t, c_int],
    "rcabeam_gram_f64_workspace_bytes": [c_size_t, c_size_t],
    "rcabeam_temporal_gram_f64_device": [c_void_p, c_size_t, c_size_t, c_void_p, c_void_p, c_size_t, c_int],
    "rcabeam_temporal_gram_f64_noalloc_device": [c_void_p, c_size_t, c_size_t, c_void_p, c_int],
    "rcabeam_simulate_lds_bytes": [c_size_t, c_int],
    "rcabeam_simulate_channels_device": [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_int, c_void_p, c_void_p,
                                         c_size_t, c_void_p, c_void_p, c_void_p] + [c_size_t] * 4
    + [c_int] + [c_float] * 6 + [c_ulonglong],
    "rcabeam_gram_hermitian_device": [c_void_p, c_size_t, c_double, c_void_p],
    "rcabeam_casorati_colmajor_device": [c_void_p, c_int, c_size_t, c_size_t, c_void_p, c_int, c_size_t, c_size_t],
    "rcabeam_svd_device": [c_void_p, c_int, c_size_t, c_size_t, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "rcabeam_scale_columns_device": [c_void_p, c_int, c_size_t, c_size_t, c_size_t, c_void_p, c_size_t, c_size_t],
    "rcabeam_times_square_device": [c_void_p, c_int, c_size_t, c_size_t, c_size_t, c_void_p, c_void_p, c_int,
                                    c_void_p, c_size_t],
    "rcabeam_gram_colmajor_device": [c_void_p, c_size_t, c_size_t, c_size_t, c_void_p],
    "rcabeam_rf_to_iq_device": [c_void_p, c_int] + [c_size_t] * 4 + [c_void_p, c_size_t, c_size_t] + [c_float] * 3
    + [c_void_p, c_size_t, c_int],
    "rcabeam_projector_device": [c_void_p, c_size_t, c_size_t, c_size_t, c_void_p],
    "rcabeam_eigh_device": [c_void_p, c_void_p, c_void_p, c_void_p, c_size_t],
    "rcabeam_filtered_power_device": [c_void_p, c_int, c_size_t, c_size_t, c_void_p, c_size_t, c_int, c_void_p],
    "rcabeam_temporal_filter_device": [c_void_p, c_int, c_size_t, c_size_t, c_void_p, c_size_t, c_int, c_void_p],
    "rcabeam_xdoppler_mean_device": [c_void_p, c_void_p] + [c_size_t] * 7,
    "rcabeam_angle_accumulator_fields": [c_int, c_int],
    "rcabeam_angle_accumulate_device": [c_void_p] + [c_size_t] * 4 + [c_int] * 3 + [c_void_p],
    "rcabeam_angle_combine_finish_device": [c_void_p, c_size_t, c_size_t, c_size_t, c_int, c_int, c_void_p, c_void_p,
                                            c_void_p],
    "rcabeam_dmas_accumulator_fields": [],
    "rcabeam_dmas_accumulate_device": [c_void_p] + [c_size_t] * 4 + [c_int, c_int, c_void_p],
    "rcabeam_dmas_finish_device": [c_void_p] * 4 + [c_size_t] * 3 + [c_void_p] * 2 + [c_size_t] * 3
    + [c_float, c_void_p],
    "rcabeam_two_stage_planes_device": [c_void_p] * 10 + [c_int] * 4 + [c_size_t] * 10 + [c_float] * 4 + [c_int],
    "rcabeam_two_stage_spline_device": [c_void_p] * 4 + [c_size_t] * 5,
    "rcabeam_two_stage_expand_device": [c_void_p] * 7 + [c_int] + [c_size_t] * 11 + [c_float] * 2 + [c_void_p] * 4
    + [c_size_t] + [c_void_p] * 3,
    "rcabeam_two_stage_stsw_finish_device": [c_void_p] * 3 + [c_size_t] * 2 + [c_void_p],
    "rcabeam_two_stage_buffer_elements": [c_void_p],
    "rcabeam_two_stage_reconstruct_device": [c_void_p] * 5,
    "rcabeam_hybrid_planes_device": [c_void_p] * 4 + [c_size_t] * 6,
    "rcabeam_hybrid_tables_device": [c_void_p] * 3 + [c_size_t] * 7 + [c_float],
    "rcabeam_hybrid_finish_device": [c_void_p] * 11 + [c_size_t] * 10 + [c_float, c_int],
}
_RESTYPE = {
    "rcabeam_status_string": c_char_p,
    "rcabeam_gram_workspace_bytes": c_size_t,
    "rcabeam_gram_real_elements": c_size_t,
    "rcabeam_gram_f64_workspace_bytes": c_size_t,
    "rcabeam_opw_adjoint_workspace_bytes": c_size_t,
    "rcabeam_simulate_lds_bytes": c_size_t,
    "rcabeam_two_stage_buffer_elements": c_size_t,
    "rcabeam_angle_accumulator_fields": c_size_t,
    "rcabeam_dmas_accumulator_fields": c_size_t,
}


class TwoStagePlan(ctypes.Structure):
    """`rcabeam_two_stage_plan` (include/rcabeam_mi355x.h); pointers are plain addresses."""

    _fields_ = [(name, c_size_t) for name in ("ns", "nc", "na", "nt", "nx", "nz", "ny", "npz", "n_planes", "chunk",
                                               "n_bands")]
    _fields_ += [(name, c_void_p) for name in (
        "band_depth_start", "band_depth_count", "band_knot_first", "band_knot_count", "band_radius", "knots_host",
        "x", "z", "y", "x_el", "y_el", "angles", "starts", "slopes", "knots", "x_centers", "x_plane", "x_offset",
        "y_centers", "y_plane", "y_offset", "x_same", "y_same", "spline_upper", "spline_inverse", "x_phase",
        "z_phase", "y_phase")]
    _fields_ += [(name, c_float) for name in ("c", "fs", "fd", "dz")]
    _fields_ += [("baseband", c_int), ("stsw", c_int)]

# Symbols declared in include/*.h that the library must export.
EXPORTED = tuple(_SIGNATURES)

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def library_path() -> Path:
    return LIBRARY_PATH


def load(require_gpu: bool = True) -> ctypes.CDLL:
    """Load the shared library (once).  With `require_gpu`, also require a HIP device."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIBRARY_PATH.exists():
                raise RuntimeError(
                    f"rcabeam MI355X library not built: {LIBRARY_PATH} is missing "
                    "(run `python -m rcabeam_amd.build`)"
                )
            lib = ctypes.CDLL(str(LIBRARY_PATH))
            for name, argtypes in _SIGNATURES.items():
                fn = getattr(lib, name, None)
                if fn is None:
                    continue
                fn.argtypes = argtypes
                fn.restype = _RESTYPE.get(name, c_int)
            _lib = lib
    if require_gpu:
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("rcabeam MI355X kernels need a HIP GPU; none is visible")
    return _lib


def status_string(status: int) -> str:
    return load(require_gpu=False).rcabeam_status_string(status).decode()


def check(status: int, what: str) -> None:
    """Raise for a non-success status, as the reference bindings do."""
    if status == SUCCESS:
        return
    message = f"{what}: {status_string(status)}"
    if status == ERROR_ARGUMENT:
        raise ValueError(message)
    raise RuntimeError(message)


def call(name: str, *args) -> None:
    """Call an entry point by name and check its status."""
    check(getattr(load(), name)(*args), name)
