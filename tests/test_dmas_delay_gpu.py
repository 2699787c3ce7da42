"""Hybrid DMAS-CCF-ACF (rcabeam_amd.dmas_delay, reference examples/_support/dmas_delay.py) on the GPU.

* The reference's own check (tests/test_dmas_delay_benchmark.py:11-52): on the broadside case the
  hybrid equals exact DMAS-CCF-ACF (float64 oracle) within rtol 3e-4, atol 2e-6 of the maximum,
  linear and cubic tables; a common gain of 2 doubles the power.
* Steered point scene over several depth slabs, against the float64 restatement
  (oracle/dmas_delay.py) fed the same float32 geometry: relative L2 error below 2e-5 (float32
  table sums and DMAS cancellation |sum of signed roots|^2 - sum of magnitudes; measured 0.8e-6 to
  4e-6).
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import dmas_delay as oracle_hybrid
from oracle import scene
from rcabeam_amd.dmas_delay import prepare_hybrid

pytestmark = pytest.mark.gpu

C, FS, FD = 1540.0, 20e6, 15e6


def _broadside():
    from test_dmas_delay import broadside

    return broadside()


@pytest.mark.parametrize("cubic", [False, True])
def test_broadside_equals_exact_dmas(cubic):
    from test_dmas_delay import exact

    raw, angles, grid, x_el, y_el = _broadside()
    print(f"case: hybrid broadside cubic={cubic}", flush=True)
    geom = rb.RCAGeometry(x_el, y_el, c=C, fs=FS, f_demod=FD, fnumber=1.0)
    expected = exact(raw, angles, grid, x_el, y_el)
    reconstruct = prepare_hybrid(raw[0].shape, grid, geom, angles, 0.75, 3, cubic=cubic)
    got = reconstruct(raw)
    assert got.shape == expected.shape and got.dtype == np.float32
    np.testing.assert_allclose(got, expected, rtol=3e-4, atol=2e-6 * expected.max())
    np.testing.assert_allclose(reconstruct([2 * v for v in raw]), 2 * expected, rtol=3e-4,
                               atol=4e-6 * expected.max())


@pytest.mark.parametrize("fnumber,cubic,step,slab", [(1.0, True, 0.5, 4), (1.0, False, 0.75, 3), (None, True, 0.5, 5)])
def test_steered_scene_matches_oracle(fnumber, cubic, step, slab):
    acq = scene.point_scene(elements=16, samples=320, frames=70, grid_shape=(6, 9, 5), angles_deg=(-6.0, 0.0, 4.0),
                            t_start=oracle_hybrid.T0, fnumber=fnumber)
    print(f"case: hybrid steered fnumber={fnumber} cubic={cubic} step={step} slab={slab}", flush=True)
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=fnumber)
    expected = oracle_hybrid.hybrid_dmas(acq.rc, acq.cr, acq.angles, acq.grid, acq.x_el, acq.y_el, acq.c, acq.fs,
                                         acq.fd, fnumber, step=step, slab=slab, cubic=cubic)
    got = prepare_hybrid(acq.rc.shape, acq.grid, geom, acq.angles, step, slab, cubic=cubic)([acq.rc, acq.cr])
    err = rel_l2(got.ravel(), expected.ravel())
    print(f"rel L2 {err:.2e}, max {expected.max():.3e}", flush=True)
    assert expected.max() > 0 and np.isfinite(got).all()
    assert err < 2e-5


def test_device_tensors_stay_on_device():
    import torch

    raw, angles, grid, x_el, y_el = _broadside()
    geom = rb.RCAGeometry(x_el, y_el, c=C, fs=FS, f_demod=FD, fnumber=1.0)
    reconstruct = prepare_hybrid(raw[0].shape, grid, geom, angles, 0.75, 3)
    out = reconstruct([torch.from_numpy(v).cuda() for v in raw])
    assert isinstance(out, torch.Tensor) and out.is_cuda
    np.testing.assert_allclose(out.cpu().numpy(), reconstruct(raw), rtol=0, atol=0)
