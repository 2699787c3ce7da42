"""Two-stage OPW approximation: oracle pinning and host planning (CPU only).

The oracle (`oracle/two_stage.py`) is pinned against scipy's not-a-knot
spline, the conventional delay-and-sum oracle where the approximation is exact,
slow-time phase invariants and the aperture cases of the reference's own tests
(reference tests/test_two_stage.py).  The package's host plan
(`rcabeam_amd.two_stage`) is checked against the oracle's plan.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import two_stage as ots
from oracle.das import opw_iq, scan_coords
from oracle.scene import simulate_point

FS, FD, C = 20e6, 15e6, 1540.0


def _sim(x_el, y_el, angles, point, nsamp, t0, config):
    return simulate_point(x_el, y_el, C, FS, FD, angles, point, nsamp, t0, config)


def _case(frames=3):
    """reference tests/test_two_stage.py:10-18."""
    el = np.array([-0.1e-3, 0, 0.1e-3])
    angles = np.array([-0.12, 0.03, 0.1])
    starts = np.array([1.9e-6, 2e-6, 2.1e-6])
    grid = (np.array([-0.12e-3, 0.03e-3]), np.linspace(7.9e-3, 8.1e-3, 5), np.array([-0.2e-3, 0, 0.1e-3]))
    phase = np.exp(0.2j * np.arange(frames)).astype(np.complex64)
    raw = [_sim(el, el, angles, (0, 8e-3, 0), 368, 2e-6, cfg)[..., None] * phase for cfg in ("RC", "CR")]
    return raw[0], raw[1], angles, starts, grid, el, el, 1.0


def _conventional(rc, cr, angles, starts, grid, x_el, y_el, fnumber):
    shape = tuple(len(g) for g in grid)
    iq = opw_iq(rc, cr, scan_coords(grid), x_el, y_el, angles, starts, C, FS, FD, fnumber)
    return iq.reshape(*shape, -1)


def _two_stage(rc, cr, angles, starts, grid, x_el, y_el, fnumber, **options):
    return ots.two_stage_opw(rc, cr, angles, starts, grid, x_el, y_el, C, FS, FD, fnumber, **options)


@pytest.mark.parametrize("n", [4, 5, 6, 31])
def test_oracle_spline_matches_scipy(n):
    """The banded not-a-knot solve equals scipy's CubicSpline (reference test_two_stage.py:34-42)."""
    from scipy.interpolate import CubicSpline

    rng = np.random.default_rng(42)
    p = rng.normal(size=(2, n, 3, 4)) + 1j * rng.normal(size=(2, n, 3, 4))
    expected = CubicSpline(np.arange(n), p, axis=1)(np.arange(n), 2)
    np.testing.assert_allclose(ots.second_derivatives(p), expected, rtol=1e-12, atol=1e-12)


def test_oracle_spline_reproduces_cubics():
    """A not-a-knot spline through samples of a cubic is that cubic: M = p'' exactly."""
    t = np.arange(9.0)
    p = (0.3 * t**3 - t**2 + 2 * t - 1) + 1j * (-0.1 * t**3 + 0.5 * t)
    np.testing.assert_allclose(ots.second_derivatives(p[None], axis=1)[0], (1.8 * t - 2) + 1j * (-0.6 * t), atol=1e-10)


def test_oracle_zero_steering_matches_conventional():
    """At normal incidence the elevation mapping is exact (reference test_two_stage.py:45-56)."""
    _, _, _, _, grid, x_el, y_el, fn = _case()
    angles = np.zeros(3)
    phase = np.exp(0.2j * np.arange(3)).astype(np.complex64)
    rc, cr = [_sim(x_el, y_el, angles, (0, 8e-3, 0), 368, 2e-6, cfg)[..., None] * phase for cfg in ("RC", "CR")]
    expected = _conventional(rc, cr, angles, 2e-6, grid, x_el, y_el, fn)
    actual = _two_stage(rc, cr, angles, 2e-6, grid, x_el, y_el, fn, samples_per_wavelength=32)
    np.testing.assert_allclose(actual, expected, rtol=2e-3, atol=2e-3)


def test_oracle_closest_element_path_and_slow_time_phase():
    """One element under the voxel: the mapping is exact up to the spline (reference test_two_stage.py:59-71)."""
    el = np.array([0.0])
    angles = np.array([0.14])
    grid = (np.array([0.0]), np.linspace(7.85e-3, 8.15e-3, 21), np.array([-1e-3, 0, 1e-3]))
    phase = np.exp(0.4j * np.arange(3)).astype(np.complex64)
    rc = _sim(el, el, angles, (0, 8e-3, 0.7e-3), 368, 2e-6, "RC")[..., None] * phase
    cr = np.zeros_like(rc)
    expected = _conventional(rc, cr, angles, 2e-6, grid, el, el, None)
    actual = _two_stage(rc, cr, angles, 2e-6, grid, el, el, None, samples_per_wavelength=32)
    assert np.linalg.norm(actual - expected) / np.linalg.norm(expected) < 0.003
    np.testing.assert_allclose(actual[..., 1], actual[..., 0] * phase[1], rtol=1e-6, atol=1e-6)


def test_oracle_moving_target_phase():
    """Axial motion keeps the lag-one phase of conventional DAS (reference test_two_stage.py:167-183)."""
    el = np.array([0.0])
    angles = np.array([0.14])
    grid = (np.array([0.0]), np.linspace(7.9e-3, 8.1e-3, 21), np.array([-1e-3, 0.7e-3, 1e-3]))
    rc = np.stack([_sim(el, el, angles, (0.0, 8e-3 + t * 8e-6, 0.7e-3), 368, 2e-6, "RC") for t in range(5)], axis=-1)
    cr = np.zeros_like(rc)
    expected = _conventional(rc, cr, angles, 2e-6, grid, el, el, None)
    actual = _two_stage(rc, cr, angles, 2e-6, grid, el, el, None, samples_per_wavelength=32)
    assert np.linalg.norm(actual - expected) / np.linalg.norm(expected) < 0.005
    old = np.sum(expected[0, 10, 1, :-1] * expected[0, 10, 1, 1:].conj())
    new = np.sum(actual[0, 10, 1, :-1] * actual[0, 10, 1, 1:].conj())
    assert abs(np.angle(new * old.conj())) < 0.003


@pytest.mark.filterwarnings("ignore:Axial undersampling")
@pytest.mark.parametrize("baseband,n_planes,spw,atol", [(False, 1, 8, 0.05), (True, 1, 4, 0.002), (True, 2, 4, 0.002)])
def test_oracle_aperture_uses_output_depth(baseband, n_planes, spw, atol):
    """The receive mask is taken at the band's output depth (reference test_two_stage.py:192-224)."""
    x_el, y_el = np.array([0.00205]), np.zeros(1)
    angles = np.array([0.14])
    grid = (np.zeros(1), np.array([0.004, 0.0042, 0.004]), np.array([-0.004, 0.004]))
    rc = np.ones((256, 1, 1, 3), np.complex64)
    cr = np.zeros_like(rc)
    out = _two_stage(rc, cr, angles, 0, grid, x_el, y_el, 1.0, samples_per_wavelength=spw, n_planes=n_planes,
                     baseband=baseband)
    np.testing.assert_array_equal(out[:, [0, 2]], 0)
    np.testing.assert_allclose(np.abs(out[:, 1]), 1, atol=atol)


def test_oracle_per_angle_parts_are_linear_in_each_angle():
    """Each angle's part equals the result of that angle's data alone (reference test_two_stage.py:241-251)."""
    rng = np.random.default_rng(21)
    el = np.array([-0.0003, 0.0001])
    angles = np.array([-0.1, 0.03, 0.13])
    grid = (np.array([-0.0005, 0.0002]), np.array([0.003, 0.004]), np.array([-0.0007, 0.0003]))
    shape = (96, 2, 3, 4)
    raw = [(rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64) for _ in range(2)]
    args = (angles, 0, grid, el, el[::-1], 1540.0, 12e6, 5e6, 1.0)
    opts = dict(samples_per_wavelength=4, baseband=True)
    r, q = ots.two_stage_per_angle(*raw, *args, **opts)
    for cfg in (0, 1):
        for m in range(3):
            parts = [np.zeros_like(v) for v in raw]
            parts[cfg][:, :, m] = raw[cfg][:, :, m]
            alone = ots.two_stage_opw(*parts, *args, **opts)
            np.testing.assert_allclose(alone, (r if cfg == 0 else q)[..., m, :], rtol=1e-12, atol=1e-12)


def test_oracle_stsw_zero_input_and_weights_range():
    rc, cr, angles, starts, grid, x_el, y_el, fn = _case(8)
    w, p = ots.two_stage_stsw(rc, cr, angles, starts, grid, x_el, y_el, C, FS, FD, fn)
    assert w.max() == pytest.approx(1.0) and (w >= 0).all() and (p >= 0).all()
    z = np.zeros_like(rc)
    w0, p0 = ots.two_stage_stsw(z, z, angles, starts, grid, x_el, y_el, C, FS, FD, fn)
    np.testing.assert_array_equal(w0, 0)
    np.testing.assert_array_equal(p0, 0)


def test_oracle_plane_depths_are_padded():
    """Reference depths cover every mapped depth with two steps to spare (reference test_two_stage.py:137-149)."""
    rc, _, angles, starts, grid, x_el, y_el, fn = _case()
    p = ots.plan(rc.shape, angles, starts, grid, x_el, y_el, C, FS, FD, fn, samples_per_wavelength=8, n_planes=1)
    z = np.asarray(grid[1], np.float32)
    assert p["plane_z"].min() < z.min() + p["shifts"].min() - p["dz"]
    assert p["plane_z"].max() > z.max() + p["shifts"].max() + p["dz"]


@pytest.mark.parametrize("density", [2, 4])
def test_oracle_undersampling_warns(density):
    rc, _, angles, starts, grid, x_el, y_el, fn = _case(1)
    with pytest.warns(UserWarning, match="undersampling"):
        ots.plan(rc.shape, angles, starts, grid, x_el, y_el, C, FS, FD, fn, samples_per_wavelength=density, n_planes=1)


def test_oracle_unchanged_apertures_from_channel_sets():
    """reference tests/test_two_stage.py:102-109."""
    lateral = np.array([0, -2, 1], np.float32)
    elements = np.array([2, 0, -1, 2], np.float32)
    depths = np.array([1, 2, 4, 2, 1], np.float32)
    flags = ots.unchanged_apertures(lateral, elements, depths, 1.0)
    active = abs(lateral[:, None] - elements) <= depths[:, None, None] / 2
    assert not flags[0].any()
    np.testing.assert_array_equal(flags[1:], (active[1:] == active[:-1]).all(axis=-1))


def _package_plan(rc_shape, angles, starts, grid, x_el, y_el, fnumber, **options):
    from rcabeam_amd.sim import RCAGeometry
    from rcabeam_amd.two_stage import _Plan

    geom = RCAGeometry(np.asarray(x_el, np.float32), np.asarray(y_el, np.float32), c=C, fs=FS, f_demod=FD,
                       fnumber=fnumber)
    return _Plan(rc_shape, angles, starts, grid, geom, options.get("samples_per_wavelength", 8),
                 options.get("n_planes", 1), options.get("baseband", False))


@pytest.mark.parametrize("n_planes,spw,fnumber", [(1, 8, 1.0), (2, 4, 1.0), (3, 8, None), (1, 4, 0.7)])
def test_package_plan_matches_oracle(n_planes, spw, fnumber):
    """Host planning (slabs, reference depths, bands, unchanged-aperture flags) equals the oracle's."""
    el = ((np.arange(24) - 11.5) * 200e-6).astype(np.float32)
    angles = np.deg2rad(np.linspace(-10, 10, 5))
    grid = (np.linspace(-2.4e-3, 2.4e-3, 13), np.linspace(5e-3, 9e-3, 33), np.linspace(-2e-3, 2.6e-3, 11))
    shape = (512, 24, 5, 4)
    ours = _package_plan(shape, angles, 0, grid, el, el[::-1], fnumber, n_planes=n_planes, samples_per_wavelength=spw,
                         baseband=True)
    ref = ots.plan(shape, angles, 0, grid, el, el[::-1], C, FS, FD, fnumber, samples_per_wavelength=spw,
                   n_planes=n_planes, baseband=True)
    np.testing.assert_array_equal(ours.plane_z, ref["plane_z"])
    np.testing.assert_array_equal(ours.slopes, ref["slopes"])
    for mine, theirs in zip(ours.layouts, ref["layouts"]):
        for a, b in zip(mine, theirs):
            np.testing.assert_array_equal(a, b)
    bands = ots.aperture_bands(ref)
    assert [(b[0], b[1], b[2], b[3]) for b in ours.bands] == [
        (d.start, d.stop - d.start, k.start, k.stop - k.start) for d, k in bands]
    if fnumber is not None:
        assert len(bands) > 1
        depths = np.array([ref["axes"][1][d.start] for d, _ in bands], np.float32)
        for axis, elements in ((0, 0), (2, 1)):
            np.testing.assert_array_equal(
                ours.unchanged(ours.axes[axis], ours.elements[elements]),
                ots.unchanged_apertures(ref["axes"][axis], ref["elements"][elements], depths, fnumber))


@pytest.mark.parametrize("density,chunk", [(0, 32), (np.nan, 32), (8, 0)])
def test_invalid_sampling(density, chunk):
    """reference tests/test_two_stage.py:129-134 (rejected before any device work)."""
    from rcabeam_amd.sim import RCAGeometry
    from rcabeam_amd.two_stage import TwoStageOpwWorkspace

    rc, _, angles, starts, grid, x_el, y_el, fn = _case()
    geom = RCAGeometry(x_el, y_el, c=C, fs=FS, f_demod=FD, fnumber=fn)
    with pytest.raises(ValueError):
        TwoStageOpwWorkspace(rc.shape, angles, starts, grid, geom, samples_per_wavelength=density, frame_chunk=chunk)


@pytest.mark.parametrize("options", [{"k": 3}, {"iq_storage": "float16"}])
def test_baseband_stsw_requires_two_subsets_and_float32(options):
    """reference tests/test_two_stage.py:279-284."""
    from rcabeam_amd import ensemble_pd_from_channels
    from rcabeam_amd.sim import RCAGeometry

    rc, cr, angles, starts, grid, x_el, y_el, fn = _case()
    geom = RCAGeometry(x_el, y_el, c=C, fs=FS, f_demod=FD, fnumber=fn)
    with pytest.raises(ValueError, match="k=2 and float32"):
        ensemble_pd_from_channels(rc, cr, angles, starts, grid, geom, method="st_sw_baseband", **options)


def test_plan_struct_matches_c_header(tmp_path):
    """The ctypes `TwoStagePlan` has the field offsets and size of `rcabeam_two_stage_plan`."""
    import ctypes
    import shutil
    import subprocess
    from pathlib import Path

    from rcabeam_amd._lib import TwoStagePlan

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    include = Path(__file__).resolve().parents[1] / "include"
    names = [f[0] for f in TwoStagePlan._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'  printf("{n} %zu\\n", offsetof(rcabeam_two_stage_plan, {n}));' for n in names)
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"rcabeam_mi355x.h\"\nint main(void) {\n"
                   f"{body}\n  printf(\"sizeof %zu\\n\", sizeof(rcabeam_two_stage_plan));\n  return 0;\n}}\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-I", str(include), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                       text=True).stdout.splitlines())
    for n in names:
        assert int(out[n]) == getattr(TwoStagePlan, n).offset, n
    assert int(out["sizeof"]) == ctypes.sizeof(TwoStagePlan)


def test_per_frame_functions_refuse_a_cpu_path():
    """This package has no CPU path: `use_cuda=False` is refused before any device work."""
    import rcabeam_amd as rb

    rc, cr, angles, starts, grid, x_el, y_el, fn = _case(1)
    geom = rb.RCAGeometry(x_el, y_el, c=C, fs=FS, f_demod=FD, fnumber=fn)
    for f in (rb.opw_pd_from_channels, rb.fast_pd_from_channels, rb.dmas_ccf_acf_from_channels):
        with pytest.raises(ValueError, match="no CPU path"):
            f(rc[..., 0], cr[..., 0], angles, starts, grid, geom, use_cuda=False)
    with pytest.raises(ValueError, match="no CPU path"):
        rb.opw(np.zeros((2, 3), np.complex64), use_cuda=False)


GOLDEN_TS = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "two_stage.npz")
GOLDEN_KEYS = [f"{fn}_p{p}_{b}" for fn in ("f8", "full") for p in (1, 2) for b in ("rf", "bb")]


def golden_two_stage_case(key):
    """Inputs of one `tests/golden/two_stage.npz` case (written by the reference `two_stage_opw_numpy`):
    (rc, cr, angles, starts, grid, x_el, y_el, fnumber, options, expected)."""
    g = GOLDEN_TS
    fn = 8.0 if key.startswith("f8") else None
    options = {"samples_per_wavelength": 8, "n_planes": int(key.split("_p")[1][0]), "baseband": key.endswith("bb")}
    grid = (g["grid_x"], g["grid_z"], g["grid_y"])
    return g["rc"], g["cr"], g["angles"], g["starts"], grid, g["x_el"], g["y_el"], fn, options, g[key]


@pytest.mark.parametrize("key", GOLDEN_KEYS)
def test_oracle_matches_reference_golden_steered(key):
    """The steered slab mapping z + o tan(theta / 2), the aperture bands (f-number 8 cuts the element row),
    one and two planes, RF and baseband planes, against the reference's own output.

    With the reference's float32 plane delays (its plane grids are float32 arrays, so its NumPy delay
    runs in float32) the restatement agrees to the fixture's complex64 rounding: rel L2 < 1e-6.  The
    float64 delays the GPU kernels use differ from the reference by that float32 rounding alone: tau
    ~5 us carries ~2^-24 relative (3e-13 s), i.e. ~3e-5 rad of carrier phase at 15 MHz and ~1e-5 samples
    at 20 MHz per tap, a few 1e-5 relative on the result (bar 1e-4)."""
    rc, cr, angles, starts, grid, x_el, y_el, fn, options, expected = golden_two_stage_case(key)
    got32 = ots.two_stage_opw(rc, cr, angles, starts, grid, x_el, y_el, C, FS, FD, fn, float32_delays=True,
                              **options)
    err32 = np.linalg.norm(got32 - expected) / np.linalg.norm(expected)
    got = ots.two_stage_opw(rc, cr, angles, starts, grid, x_el, y_el, C, FS, FD, fn, **options)
    err = np.linalg.norm(got - expected) / np.linalg.norm(expected)
    assert err32 < 1e-6, err32
    assert err < 1e-4, err


def test_golden_detects_a_dropped_slab_mapping():
    """A planted mistake in the oracle (the elevation-offset term o tan(theta / 2) of the steered mapping
    dropped, i.e. every voxel read at its own depth) fails the golden comparison: the fixture pins the
    steered mapping, which the zero-steering and closest-element checks leave open."""
    rc, cr, angles, starts, grid, x_el, y_el, fn, options, expected = golden_two_stage_case("full_p2_rf")
    original = ots.plan

    def wrong_plan(*args, **kwargs):
        p = original(*args, **kwargs)
        p["slopes"] = np.zeros_like(p["slopes"])
        return p

    try:
        ots.plan = wrong_plan
        got = ots.two_stage_opw(rc, cr, angles, starts, grid, x_el, y_el, C, FS, FD, fn, float32_delays=True,
                                **options)
    finally:
        ots.plan = original
    assert np.linalg.norm(got - expected) / np.linalg.norm(expected) > 1e-3
