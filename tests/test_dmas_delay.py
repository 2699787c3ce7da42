"""Hybrid DMAS-CCF-ACF oracle (reference examples/_support/dmas_delay.py) pinned on the CPU.

The reference's own check (tests/test_dmas_delay_benchmark.py:11-52): on a broadside case with
IQ affine in the sample index, table interpolation is exact, so the hybrid equals exact
DMAS-CCF-ACF within rtol 3e-4 (atol 2e-6 of the maximum); the power scales linearly with a common
positive channel gain.
"""

import numpy as np
import pytest

from oracle import combiners, das, dmas_delay

C, FS, FD = 1540.0, 20e6, 15e6


def broadside():
    """Four elements, three zero angles, depths on carrier-wavelength knots, 35 frames of affine IQ
    (tests/test_dmas_delay_benchmark.py:15-27)."""
    elements = np.array([-0.4e-3, 0.1e-3, 1.8e-3, 4.9e-3])
    angles = np.zeros(3)
    z = 8e-3 + np.arange(5) * C / FD
    grid = (np.array([-0.2e-3, 0.4e-3]), z, np.array([-0.8e-3, -0.1e-3, 0.5e-3]))
    rng = np.random.default_rng(42)
    raw = []
    for _ in range(2):
        coefficients = rng.standard_normal((2, 4, 3, 35)) + 1j * rng.standard_normal((2, 4, 3, 35))
        raw.append(np.complex64(coefficients[0] + np.arange(260)[:, None, None, None] * coefficients[1] / 200))
    return raw, angles, grid, elements, elements[::-1].copy()


def exact(raw, angles, grid, x_el, y_el, fnumber=1.0):
    rc_res, cr_res = das.delay_apertures(raw[0], raw[1], das.scan_coords(grid), x_el, y_el, angles, dmas_delay.T0,
                                         C, FS, FD, fnumber, channels_out=True)
    return combiners.ensemble_power("dmas", rc_res, cr_res)[0].reshape(tuple(len(g) for g in grid))


@pytest.mark.parametrize("cubic", [False, True])
def test_broadside_equals_exact_dmas(cubic):
    raw, angles, grid, x_el, y_el = broadside()
    expected = exact(raw, angles, grid, x_el, y_el)
    got = dmas_delay.hybrid_dmas(raw[0], raw[1], angles, grid, x_el, y_el, C, FS, FD, 1.0, step=0.75, slab=3,
                                 cubic=cubic)
    assert expected.max() > 0
    np.testing.assert_allclose(got, expected, rtol=3e-4, atol=2e-6 * expected.max())


def test_common_gain_scales_power():
    raw, angles, grid, x_el, y_el = broadside()
    one = dmas_delay.hybrid_dmas(raw[0], raw[1], angles, grid, x_el, y_el, C, FS, FD, 1.0, step=0.5, slab=2,
                                 cubic=True)
    two = dmas_delay.hybrid_dmas(2 * raw[0], 2 * raw[1], angles, grid, x_el, y_el, C, FS, FD, 1.0, step=0.5,
                                 slab=2, cubic=True)
    np.testing.assert_allclose(two, 2 * one, rtol=1e-9, atol=1e-12 * one.max())


def test_planted_mistakes_break_broadside():
    """A conjugated anchor phase or swapped apertures move the oracle off exact DMAS."""
    raw, angles, grid, x_el, y_el = broadside()
    expected = exact(raw, angles, grid, x_el, y_el)
    swapped = dmas_delay.hybrid_dmas(raw[1], raw[0], angles, grid, x_el, y_el, C, FS, FD, 1.0, step=0.75, slab=3)
    assert not np.allclose(swapped, expected, rtol=3e-4, atol=2e-6 * expected.max())
    original = dmas_delay._rotate
    try:
        dmas_delay._rotate = lambda v, g: original(v, g * np.array([1, 1, -1, 1], np.float32))
        conj = dmas_delay.hybrid_dmas(raw[0], raw[1], angles, grid, x_el, y_el, C, FS, FD, 1.0, step=0.75, slab=3)
    finally:
        dmas_delay._rotate = original
    assert not np.allclose(conj, expected, rtol=3e-4, atol=2e-6 * expected.max())


def test_decreasing_depths_rejected():
    raw, angles, grid, x_el, y_el = broadside()
    with pytest.raises(ValueError, match="nondecreasing depth axis"):
        dmas_delay.hybrid_geometry(raw[0].shape, (grid[0], grid[1][::-1], grid[2]), x_el, y_el, angles, C, FS, FD,
                                   1.0, 0.5, 2)


def test_product_rejects_decreasing_depths_and_bad_tables():
    """prepare_hybrid validates before touching the GPU (dmas_delay.py:63-65)."""
    import rcabeam_amd as rb
    from rcabeam_amd.dmas_delay import prepare_hybrid

    raw, angles, grid, x_el, y_el = broadside()
    geom = rb.RCAGeometry(x_el, y_el, c=C, fs=FS, f_demod=FD, fnumber=1.0)
    with pytest.raises(ValueError, match="nondecreasing depth axis"):
        prepare_hybrid(raw[0].shape, (grid[0], grid[1][::-1], grid[2]), geom, angles, 0.5, 2)
    with pytest.raises(ValueError, match="positive step and slab"):
        prepare_hybrid(raw[0].shape, grid, geom, angles, 0.0, 2)
