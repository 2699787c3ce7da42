"""Single-frame fused power volumes and per-angle combiners on the GPU against the float64 oracle.

Bar: relative L2 error 1e-5 (float32 arithmetic on complex64 data), as the other parity tests.
Cases follow reference tests/test_opw.py, test_xdoppler.py, test_fmas.py and test_dmas.py.
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import combiners, das, scene

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _frame(fnumber, seed):
    grid = (np.linspace(-0.002, 0.002, 7), np.linspace(0.002, 0.008, 5), np.linspace(-0.0015, 0.002, 6))
    acq = scene.random_acquisition(samples=24, channels=6, angles=4, frames=1, seed=seed, grid=grid,
                                   fnumber=fnumber)
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    starts = np.array([1e-6, 2e-6, 1e-6, 3e-6], np.float32)
    rc_res, cr_res = das.delay_apertures(acq.rc, acq.cr, das.scan_coords(grid), acq.x_el, acq.y_el, acq.angles,
                                         starts, acq.c, acq.fs, acq.fd, acq.fnumber, channels_out=True)
    args = (acq.rc[..., 0], acq.cr[..., 0], acq.angles, starts, grid, geom)
    return args, rc_res, cr_res


@pytest.mark.parametrize("fnumber", [None, 1.0])
def test_fused_frame_powers_match_oracle(fnumber):
    """reference fused.py:62-298 against the frame-batched oracle at one frame."""
    args, rc_res, cr_res = _frame(fnumber, 9)
    expected = {m: combiners.ensemble_power(m, rc_res, cr_res)[0] for m in ("opw", "xdoppler", "rc_fmas", "dmas")}
    got = {
        "opw": rb.opw_pd_from_channels(*args),
        "xdoppler": rb.xdoppler_pd_from_channels(*args),
        "rc_fmas": rb.rc_fmas_pd_from_channels(*args),
        "dmas": rb.dmas_ccf_acf_from_channels(*args),
    }
    for method, value in got.items():
        assert value.shape == (7, 5, 6) and value.dtype == np.float32, method
        assert rel_l2(value.ravel(), expected[method]) < TOL, method
    fast = rb.fast_pd_from_channels(*args)
    for value, method in zip(fast, ("opw", "xdoppler", "rc_fmas")):
        assert rel_l2(value.ravel(), expected[method]) < TOL, method


def test_per_angle_combiners_match_oracle():
    """opw / xdoppler_pd / rc_fmas_pd on per-angle volumes, contiguous and scattered angle sets
    (reference opw.py:24-51, xdoppler.py:65-117, fmas.py:54-89)."""
    rng = np.random.default_rng(4)
    iq = (rng.normal(size=(3, 4, 5, 6, 9)) + 1j * rng.normal(size=(3, 4, 5, 6, 9))).astype(np.complex64)
    np.testing.assert_allclose(rb.opw(iq), iq.astype(np.complex128).sum(axis=-2), rtol=1e-5, atol=1e-5)
    for rc_idx, cr_idx in ((np.arange(3), np.arange(3, 6)), (np.array([0, 4]), np.array([5, 1, 2]))):
        signal = combiners.xdoppler_signal(iq, rc_idx, cr_idx).mean(axis=-1)
        assert rel_l2(rb.xdoppler_pd(iq, rc_idx, cr_idx).ravel(), np.abs(signal).ravel()) < TOL
        assert rel_l2(rb.xdoppler_pd(iq, rc_idx, cr_idx, output="real").ravel(), signal.real.ravel()) < TOL
        assert rel_l2(rb.xdoppler_pd(iq, rc_idx, cr_idx, output="complex").ravel(), signal.ravel()) < TOL
        fmas = combiners.power_doppler(combiners.rc_fmas(iq, rc_idx, cr_idx))
        assert rel_l2(rb.rc_fmas_pd(iq, rc_idx, cr_idx).ravel(), fmas.ravel()) < TOL


def test_xdoppler_known_answer():
    """reference tests/test_xdoppler.py:35-44: RC angles sum to 1 + 2i, CR to 3 + 4i; |mean| = |(1+2i)(3-4i)| = 5*sqrt(5)."""
    iq = np.zeros((1, 4, 2), np.complex64)
    iq[0, 0] = iq[0, 1] = 0.5 + 1j
    iq[0, 2] = 3
    iq[0, 3] = 4j
    got = rb.xdoppler_pd(iq, np.arange(2), np.arange(2, 4))
    np.testing.assert_allclose(got, [abs((1 + 2j) * np.conj(3 + 4j))], rtol=1e-6)


def test_zero_input_gives_zero():
    args, _, _ = _frame(1.0, 2)
    zeros = np.zeros_like(args[0])
    for fn in (rb.opw_pd_from_channels, rb.xdoppler_pd_from_channels, rb.rc_fmas_pd_from_channels,
               rb.dmas_ccf_acf_from_channels):
        assert not fn(zeros, zeros, *args[2:]).any(), fn.__name__
