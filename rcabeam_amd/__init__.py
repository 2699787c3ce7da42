"""MI355X-native OPW row-column beamforming (drop-in for the reference `rcabeam` hot path).

Host API mirrors the reference package (src/rcabeam/__init__.py:3-49) for the
beamforming / clutter-filter / power-Doppler path.  Every numerical entry
point calls the HIP library through the C ABI (include/rcabeam.h,
include/rcabeam_mi355x.h); there is no CPU fallback.
"""

from rcabeam_amd import clutter
from rcabeam_amd.clutter_host import svd_clutter_filter, temporal_basis
from rcabeam_amd.ensemble import (
    EnsembleTiming,
    beamform_apertures,
    beamform_ensemble,
    ensemble_pd_from_channels,
    filtered_xdoppler_pd,
    opw_voxel_order,
)
from rcabeam_amd.fused import (
    dmas_ccf_acf_from_channels,
    fast_pd_from_channels,
    opw,
    opw_pd_from_channels,
    rc_fmas_pd,
    rc_fmas_pd_from_channels,
    xdoppler_pd,
    xdoppler_pd_from_channels,
)
from rcabeam_amd.sim import RCAGeometry, delay_rca_channel_data, delay_rca_channels, depth_axis

__all__ = [
    "EnsembleTiming",
    "RCAGeometry",
    "beamform_apertures",
    "beamform_ensemble",
    "clutter",
    "delay_rca_channel_data",
    "delay_rca_channels",
    "depth_axis",
    "dmas_ccf_acf_from_channels",
    "ensemble_pd_from_channels",
    "fast_pd_from_channels",
    "filtered_xdoppler_pd",
    "opw",
    "opw_pd_from_channels",
    "opw_voxel_order",
    "rc_fmas_pd",
    "rc_fmas_pd_from_channels",
    "svd_clutter_filter",
    "temporal_basis",
    "xdoppler_pd",
    "xdoppler_pd_from_channels",
]
