"""Per-frame delay entry points of the C ABI against the reference's own NumPy outputs.

`rcabeam_delay_rca_channels_device` (reference include/rcabeam.h:49-65, core.cu:192-210) and
`rcabeam_delay_rca_channel_data_device` (reference include/rcabeam.h:67-85, core.cu:244-283) through
`rb.delay_rca_channels` / `rb.delay_rca_channel_data`, frame by frame, against
`tests/golden/delay.npz`, which the reference `_delay_numpy` wrote (sim.py:344-393: both apertures,
per-angle starts, f-number 3 and the full aperture, 4 frames).

Bars: per-angle and per-channel sums within relative L2 1e-5 (float32 arithmetic on complex64 data;
the fixture itself carries the reference's complex64 rounding); `n_active` bit-exact.  The active
count uses the same float64 aperture test as the taps (|v - e| <= z / 2F), as the reference's NumPy
path does (sim.py:373-375), so counts and sums always agree with each other.
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import das, scene

pytestmark = pytest.mark.gpu

GOLDEN = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "delay.npz")


def _geometry(fnumber):
    return rb.RCAGeometry(GOLDEN["x_el"], GOLDEN["y_el"], c=1540.0, fs=20e6, f_demod=15e6, fnumber=fnumber)


@pytest.mark.parametrize("config", ["RC", "CR"])
@pytest.mark.parametrize("fn_name,fnumber", [("f3", 3.0), ("full", None)])
def test_delay_entry_points_match_reference_golden(config, fn_name, fnumber):
    """reference tests/test_cuda_api.py:50-66 and 159-171, on the reference-written fixture."""
    grid = (GOLDEN["grid_x"], GOLDEN["grid_z"], GOLDEN["grid_y"])
    key = f"{config}_{fn_name}"
    geom = _geometry(fnumber)
    iq = GOLDEN["iq"]
    print(f"case: golden delay {key}", flush=True)
    for frame in range(iq.shape[-1]):
        frame_iq = np.ascontiguousarray(iq[..., frame])
        per_angle = rb.delay_rca_channels(frame_iq, GOLDEN["angles"], GOLDEN["starts"], grid, geom, config)
        assert per_angle.shape == (3, 5, 4, 3) and per_angle.dtype == np.complex64
        assert rel_l2(per_angle, GOLDEN[f"{key}_angles"][..., frame]) < 1e-5, frame
        channel_sum, angle_sum, n_active = rb.delay_rca_channel_data(
            frame_iq, GOLDEN["angles"], GOLDEN["starts"], grid, geom, config
        )
        assert channel_sum.shape == (3, 5, 4, 5) and n_active.shape == (3, 5, 4)
        assert rel_l2(channel_sum, GOLDEN[f"{key}_channels"][..., frame]) < 1e-5, frame
        assert rel_l2(angle_sum, GOLDEN[f"{key}_angles"][..., frame]) < 1e-5, frame
        np.testing.assert_array_equal(n_active, GOLDEN[f"{key}_counts"])


@pytest.mark.parametrize("config", ["RC", "CR"])
def test_delay_entry_points_on_point_target(config):
    """The reference's own case (tests/test_cuda_api.py:50-66, 159-171): 12 (10) elements, fs = 4 f0,
    a point at 12 mm, against the float64 oracle; per-channel sums and counts included."""
    f0 = 6e6
    for n_el in (12, 10):
        el = ((np.arange(n_el) - (n_el - 1) / 2) * 0.2e-3).astype(np.float32).astype(np.float64)
        geom = rb.RCAGeometry(el, el, fs=4 * f0, f_demod=f0)
        angles = np.deg2rad(np.linspace(-4, 4, 3)).astype(np.float32).astype(np.float64)
        grid = tuple(np.asarray(a, np.float32).astype(np.float64) for a in (
            np.linspace(-0.5e-3, 0.5e-3, 4), np.linspace(11.8e-3, 12.2e-3, 4), np.linspace(-0.5e-3, 0.5e-3, 3)))
        ch = scene.simulate_point(el, el, geom.c, geom.fs, geom.f_demod, angles, (0.0, 12e-3, 0.0), 320, 8e-6, config)
        scan = das.scan_coords(grid)
        per_angle, per_channel, counts = das.delay_volumes(ch, scan, el, angles, 8e-6, geom.c, geom.fs, geom.f_demod,
                                                           geom.fnumber, config, channels_out=True)
        got = rb.delay_rca_channels(ch, angles, 8e-6, grid, geom, config)
        assert rel_l2(got.reshape(-1, 3), per_angle[..., 0]) < 1e-5
        channel_sum, angle_sum, n_active = rb.delay_rca_channel_data(ch, angles, 8e-6, grid, geom, config)
        assert rel_l2(channel_sum.reshape(-1, n_el), per_channel[..., 0]) < 1e-5
        assert rel_l2(angle_sum.reshape(-1, 3), per_angle[..., 0]) < 1e-5
        np.testing.assert_array_equal(n_active.ravel(), counts)
