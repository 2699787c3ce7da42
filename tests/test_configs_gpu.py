"""The stated workloads at their stated sizes, against the float64 oracle.

* Plain OPW + power Doppler on the 64 + 64-element acquisition (200 um pitch, IQ 256 x 64 x 3 x 16,
  angles -5/0/+5 degrees, t_start 0, five seeded scatterers plus -30 dB noise, seed 42) over the
  whole 32 x 64 x 32 grid (x, y over +/-6.4 mm, z over 5-15 mm).
* The XDoppler, RC-FMAS, DMAS-CCF-ACF and St-SW (k = 2, 3) combiners on that same acquisition and
  grid (reference kernels ensemble.cu:143-627; per-method check reference tests/test_ensemble.py:19-72).

Bar: relative L2 error 1e-5 for complex64 IQ and every power volume (float32 arithmetic on
complex64 data against the float64 restatement of the reference).
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import combiners, das, scene

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def point_case():
    """The 64 + 64-element point-scatterer acquisition and its float64 delay-and-sum (~12 s on the CPU)."""
    acq = scene.point_scene()
    assert acq.rc.shape == (256, 64, 3, 16)
    scan = das.scan_coords(acq.grid)
    res = das.delay_apertures(acq.rc, acq.cr, scan, acq.x_el, acq.y_el, acq.angles, acq.t_start, acq.c, acq.fs,
                              acq.fd, acq.fnumber, channels_out=True)
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    return acq, geom, res


def test_opw_full_grid(point_case):
    acq, geom, ((ra, _, _), (qa, _, _)) = point_case
    print("case: OPW 64+64 el, 32x64x32 grid", flush=True)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    args = (acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom)
    iq = rb.beamform_ensemble(*args)
    assert iq.shape == (32, 64, 32, 16) and iq.dtype == np.complex64
    assert rel_l2(iq.reshape(-1, 16), expected) < TOL
    pd = rb.ensemble_pd_from_channels(*args)
    assert pd.shape == (32, 64, 32) and pd.dtype == np.float32
    assert rel_l2(pd.ravel(), combiners.power_doppler(expected)) < TOL


@pytest.mark.parametrize("method,k", [("xdoppler", 2), ("rc_fmas", 2), ("dmas", 2), ("st_sw", 2), ("st_sw", 3)])
def test_combiners_full_grid(point_case, method, k):
    acq, geom, (rc_res, cr_res) = point_case
    print(f"case: {method} k={k} 64+64 el, 32x64x32 grid", flush=True)
    power, signal = combiners.ensemble_power(method, rc_res, cr_res, k=k)
    args = (acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom)
    got = rb.ensemble_pd_from_channels(*args, method=method, k=k)
    assert got.shape == (32, 64, 32) and got.dtype == np.float32
    assert rel_l2(got.ravel(), power) < TOL
    if method in ("xdoppler", "rc_fmas"):
        assert rel_l2(rb.beamform_ensemble(*args, method=method).reshape(-1, 16), signal) < TOL
