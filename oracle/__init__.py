"""CPU oracle for the OPW row-column beamforming path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package (`rcabeam_amd/`) may
import, call or link anything under `oracle/`.  Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` use it, and
only as the checker (or as the timed CPU baseline), never as the product path.

Every function restates the reference's NumPy algorithm in float64 and cites
the reference file:line it follows (paths relative to the reference repo root).
Inputs that cross the reference's C ABI as `float` (scan coordinates, element
positions, angles, start times, c, fs, f_demod, f-number) are rounded to
float32 first by `abi_round`, so the oracle computes exactly the problem the
GPU kernels receive, in float64.

Pinning (what checks each function beyond retyping its formula) is listed in
each module header; `tests/test_oracle.py` holds the checks.
"""

from oracle.das import abi_round, delay_apertures, delay_volumes, opw_iq, scan_coords
from oracle.combiners import (
    dmas_bruteforce,
    dmas_ccf_acf_core,
    ensemble_power,
    opw_numpy,
    power_doppler,
    rc_fmas,
    signed_sqrt,
    st_sw_weights,
    xdoppler_signal,
)
from oracle.clutter import filtered_power, rank_interval, svd_clutter_filter, temporal_basis

__all__ = [
    "abi_round",
    "delay_apertures",
    "delay_volumes",
    "dmas_bruteforce",
    "dmas_ccf_acf_core",
    "ensemble_power",
    "filtered_power",
    "opw_iq",
    "opw_numpy",
    "power_doppler",
    "rank_interval",
    "rc_fmas",
    "scan_coords",
    "signed_sqrt",
    "st_sw_weights",
    "svd_clutter_filter",
    "temporal_basis",
    "xdoppler_signal",
]
