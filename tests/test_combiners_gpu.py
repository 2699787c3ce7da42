"""RC-FMAS, St-SW and DMAS-CCF-ACF of a whole grid from per-angle / per-channel grid launches, against
the float64 oracle.

Bar: relative L2 error 1e-5 (float32 storage), as the ensemble parity tests.
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import combiners, das, scene

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _case(angles, frames, grid_shape, seed):
    nx, nz, ny = grid_shape
    grid = (np.linspace(-0.002, 0.002, nx), np.linspace(0.001, 0.008, nz), np.linspace(-0.0015, 0.002, ny))
    acq = scene.random_acquisition(samples=24, channels=6, angles=angles, frames=frames, seed=seed, grid=grid)
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    scan = das.scan_coords(acq.grid)
    res = das.delay_apertures(acq.rc, acq.cr, scan, acq.x_el, acq.y_el, acq.angles, acq.t_start, acq.c, acq.fs,
                              acq.fd, acq.fnumber, channels_out=True)
    return acq, geom, res


@pytest.mark.parametrize("budget", [None, 1])
def test_slabs_match_oracle(monkeypatch, budget):
    """x-slabs of 8 (and a ragged last slab of 4) give the whole-grid result."""
    from rcabeam_amd import ensemble

    acq, geom, (rc_res, cr_res) = _case(4, 24, (20, 6, 9), 11)
    if budget is not None:
        # One byte of budget: one line of x per slab.
        monkeypatch.setattr(ensemble, "_COMBINE_BYTES", budget)
    print(f"case: slabs budget={budget}", flush=True)
    args = (acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom)
    power, signal = combiners.ensemble_power("rc_fmas", rc_res, cr_res)
    assert rel_l2(rb.ensemble_pd_from_channels(*args, method="rc_fmas").ravel(), power) < TOL
    assert rel_l2(rb.beamform_ensemble(*args, method="rc_fmas").reshape(-1, 24), signal) < TOL
    for k in (2, 4):
        power, _ = combiners.ensemble_power("st_sw", rc_res, cr_res, k=k)
        assert rel_l2(rb.ensemble_pd_from_channels(*args, method="st_sw", k=k).ravel(), power) < TOL, k
    power, _ = combiners.ensemble_power("dmas", rc_res, cr_res)
    assert rel_l2(rb.ensemble_pd_from_channels(*args, method="dmas").ravel(), power) < TOL


def test_many_angles():
    """20 angles (beyond the grid kernel's 16 per launch; one angle per launch here)."""
    acq, geom, (rc_res, cr_res) = _case(20, 9, (8, 4, 8), 5)
    print("case: 20 angles", flush=True)
    args = (acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom)
    for method, k in (("rc_fmas", 2), ("st_sw", 3)):
        power, _ = combiners.ensemble_power(method, rc_res, cr_res, k=k)
        assert rel_l2(rb.ensemble_pd_from_channels(*args, method=method, k=k).ravel(), power) < TOL, method


def test_frames_beyond_grid_limit_keep_voxel_list_path():
    """210 frames (more than the grid kernel's 208) run the voxel-list kernels, still matching the oracle."""
    acq, geom, (rc_res, cr_res) = _case(4, 210, (4, 3, 4), 3)
    args = (acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom)
    for method, k in (("rc_fmas", 2), ("st_sw", 4), ("dmas", 2)):
        power, _ = combiners.ensemble_power(method, rc_res, cr_res, k=k)
        assert rel_l2(rb.ensemble_pd_from_channels(*args, method=method, k=k).ravel(), power) < TOL, method
