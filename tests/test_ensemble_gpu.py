"""GPU parity of the frame-batched ensemble path against the float64 oracle.

Tolerance: relative L2 error <= 1e-5 for complex64 IQ and for power (the
stated bar for float32 storage), on seeded inputs small enough for the oracle.
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import combiners, das, scene

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _geom(acq):
    return rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)


def _oracle(acq, channels_out=False):
    scan = das.scan_coords(acq.grid)
    return das.delay_apertures(
        acq.rc, acq.cr, scan, acq.x_el, acq.y_el, acq.angles, acq.t_start, acq.c, acq.fs, acq.fd, acq.fnumber,
        channels_out=channels_out,
    )


def test_opw_point_scene_subgrid():
    """64+64 elements, 256 samples, 3 angles, 16 frames, 5 scatterers + -30 dB noise, 16x32x16 voxels."""
    acq = scene.point_scene(grid_shape=(16, 32, 16))
    print("case: point_scene 16x32x16", flush=True)
    (ra, _, _), (qa, _, _) = _oracle(acq)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    shape = tuple(len(g) for g in acq.grid)
    iq = rb.beamform_ensemble(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, _geom(acq))
    assert iq.shape == (*shape, 16) and iq.dtype == np.complex64
    assert rel_l2(iq.reshape(-1, 16), expected) < TOL
    pd = rb.ensemble_pd_from_channels(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, _geom(acq))
    assert pd.dtype == np.float32
    assert rel_l2(pd.ravel(), combiners.power_doppler(expected)) < TOL


@pytest.mark.parametrize("frames", [1, 33, 200])
@pytest.mark.parametrize("fnumber", [None, 1.0])
def test_all_methods_match_oracle(frames, fnumber):
    """Shape of reference tests/test_ensemble.py:31-87, checked at 1e-5 relative L2."""
    print(f"case: methods frames={frames} fnumber={fnumber}", flush=True)
    acq = scene.random_acquisition(frames=frames, fnumber=fnumber)
    geom = _geom(acq)
    starts = np.array([1e-6, 2e-6, 1e-6, 3e-6], np.float32)
    acq = scene.Acquisition(acq.rc, acq.cr, acq.x_el, acq.y_el, acq.angles, starts, acq.grid, acq.c, acq.fs,
                            acq.fd, acq.fnumber)
    rc_res, cr_res = _oracle(acq, channels_out=True)
    args = (acq.rc, acq.cr, acq.angles, starts, acq.grid, geom)
    r, q = rb.beamform_apertures(*args)
    assert rel_l2(r.reshape(-1, frames), rc_res[0].sum(axis=1)) < TOL
    assert rel_l2(q.reshape(-1, frames), cr_res[0].sum(axis=1)) < TOL
    for method in ("opw", "xdoppler", "rc_fmas"):
        power, signal = combiners.ensemble_power(method, rc_res, cr_res)
        timing = {"raw_upload_seconds": -1.0}
        actual = rb.beamform_ensemble(*args, method=method, timings=timing)
        assert timing["raw_upload_seconds"] > 0
        assert rel_l2(actual.reshape(-1, frames), signal) < TOL, method
        pd = rb.ensemble_pd_from_channels(*args, method=method)
        assert rel_l2(pd.ravel(), power) < TOL, method
    power, _ = combiners.ensemble_power("dmas", rc_res, cr_res)
    assert rel_l2(rb.ensemble_pd_from_channels(*args, method="dmas").ravel(), power) < TOL
    for k in (2, 3, 4):
        power, _ = combiners.ensemble_power("st_sw", rc_res, cr_res, k=k)
        assert rel_l2(rb.ensemble_pd_from_channels(*args, method="st_sw", k=k).ravel(), power) < TOL, k
    zeros = np.zeros_like(acq.rc)
    for method in ("opw", "xdoppler", "rc_fmas", "dmas", "st_sw"):
        out = rb.ensemble_pd_from_channels(zeros, zeros, acq.angles, starts, acq.grid, geom, method=method)
        assert not out.any(), method


def test_closed_receive_window_known_answer():
    """Reference tests/test_iq_sampling.py:42-53 through the GPU per-frame entry point."""
    zero = np.zeros(1)
    geom = rb.RCAGeometry(zero, zero, c=2, fs=1, f_demod=0, fnumber=None)
    raw = np.arange(1, 4, dtype=np.complex64).reshape(3, 1, 1)
    _, angles, _ = rb.delay_rca_channel_data(raw, zero, 0, (zero, np.array([0, 1.5, 2, 2.25]), zero), geom, "RC")
    np.testing.assert_array_equal(angles.ravel(), [1, 2.5, 3, 0])
    # The same window through the frame-batched ensemble (one frame, both apertures).
    ens = raw[..., None]
    iq = rb.beamform_apertures(ens, ens, zero, 0, (zero, np.array([0, 1.5, 2, 2.25]), zero), geom)[0]
    np.testing.assert_array_equal(iq.ravel(), [1, 2.5, 3, 0])


@pytest.mark.parametrize("ny", [4097, 16385])
def test_voxel_tail_and_large_volume(ny):
    """Voxel counts that are not multiples of the workgroup tile (reference tests/test_ensemble_targets.py:68-91)."""
    el = np.array([-0.05e-3, 0.05e-3], np.float32)
    grid = (np.zeros(1, np.float32), np.array([8e-3], np.float32), np.linspace(-1e-3, 1e-3, ny).astype(np.float32))
    acq = scene.random_acquisition(samples=368, channels=2, angles=4, frames=3, grid=grid, fs=20e6, fd=15e6, seed=7)
    acq = scene.Acquisition(acq.rc, acq.cr, el, el, acq.angles, 2e-6, grid, acq.c, acq.fs, acq.fd, 1.0)
    (ra, _, _), (qa, _, _) = _oracle(acq)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    iq = rb.beamform_ensemble(acq.rc, acq.cr, acq.angles, 2e-6, grid, _geom(acq))
    assert rel_l2(iq.reshape(-1, 3), expected) < TOL


@pytest.mark.parametrize("frames", [2, 8, 64, 66, 72, 150, 193, 200, 206, 300])
@pytest.mark.parametrize("channels", [1, 17])
def test_frame_and_channel_tails(frames, channels):
    """Reference tests/test_ensemble_targets.py:94-120 (cooperative-kernel tails).  Grid kernel frame
    blocks: 150, 193 and 200 frames end in a block of single frames (none, one and all eight of them
    inside nt), 206 in a partial pair block; 300 frames take the voxel-list kernel."""
    el = np.array([0.001], np.float32) if channels == 1 else np.linspace(-0.006, 0.001, channels).astype(np.float32)
    grid = (np.zeros(1, np.float32), np.array([0.007], np.float32), np.array([-0.001, 0.0, 0.001], np.float32))
    acq = scene.random_acquisition(channels=channels, angles=3, frames=frames, grid=grid, seed=19)
    starts = np.array([1e-6, 2e-6, 3e-6], np.float32)
    acq = scene.Acquisition(acq.rc, acq.cr, el, el, np.array([-0.08, 0.02, 0.09], np.float32), starts, grid,
                            acq.c, acq.fs, acq.fd, 1.0)
    (ra, _, _), (qa, _, _) = _oracle(acq)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    geom = _geom(acq)
    iq = rb.beamform_ensemble(acq.rc, acq.cr, acq.angles, starts, grid, geom)
    assert rel_l2(iq.reshape(-1, frames), expected) < TOL
    pd = rb.ensemble_pd_from_channels(acq.rc, acq.cr, acq.angles, starts, grid, geom)
    assert rel_l2(pd.ravel(), combiners.power_doppler(expected)) < TOL


@pytest.mark.parametrize("storage,dtype", [("float16", np.float16), ("bfloat16", None)])
@pytest.mark.parametrize("frames", [1, 3, 66, 200])
def test_half_storage_matches_quantized_fp32(storage, dtype, frames):
    """Reference tests/test_opw_fp16.py:11-44: 16-bit storage equals float32 on pre-rounded input."""
    import torch

    rng = np.random.default_rng(111)
    shape = (48, 3, 4, frames)
    rc = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    cr = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    el = np.linspace(-0.001, 0.001, 3)
    geom = rb.RCAGeometry(el, el, fs=2e6, f_demod=1e6)
    angles = np.linspace(-0.1, 0.1, 4)
    grid = (np.linspace(-0.001, 0.001, 3), np.array([0.005, 0.007]), np.linspace(-0.001, 0.003, 5))

    def quantize(x):
        if dtype is not None:
            return (x.real.astype(dtype).astype(np.float32) + 1j * x.imag.astype(dtype).astype(np.float32)).astype(
                np.complex64
            )
        t = torch.from_numpy(np.ascontiguousarray(x.view(np.float32)))
        return t.to(torch.bfloat16).to(torch.float32).numpy().view(np.complex64)

    for scale in (0.0, 1.0, 1e-4):
        a, b = rc * scale, cr * scale
        expected = rb.beamform_ensemble(quantize(a), quantize(b), angles, 1e-6, grid, geom)
        actual = rb.beamform_ensemble(a, b, angles, 1e-6, grid, geom, iq_storage=storage)
        pd = rb.ensemble_pd_from_channels(a, b, angles, 1e-6, grid, geom, iq_storage=storage)
        assert actual.dtype == np.complex64 and pd.dtype == np.float32
        # FP16 samples times hi + lo FP16 coefficients (exact products); BF16 samples widened to float32.
        tol = 1e-6
        assert rel_l2(actual, expected) < tol
        assert rel_l2(pd, np.mean(np.abs(actual) ** 2, axis=-1)) < 1e-6


def test_half_storage_rejects_invalid_values_and_methods():
    """Reference tests/test_opw_fp16.py:47-60."""
    el = np.array([0.0])
    geom = rb.RCAGeometry(el, el, fs=2e6, f_demod=1e6)
    grid = (el, np.array([0.005]), el)
    angles = np.array([0.0, 0.05])
    data = np.zeros((48, 1, 2, 3), dtype=np.complex64)
    for value in (70000.0, 70000j, float("inf"), float("nan")):
        bad = data.copy()
        bad[0, 0, 0, 0] = value
        with pytest.raises(ValueError, match="FP16 IQ requires"):
            rb.ensemble_pd_from_channels(bad, data, angles, 1e-6, grid, geom, iq_storage="float16")
    with pytest.raises(ValueError, match="OPW only"):
        rb.ensemble_pd_from_channels(data, data, angles, 1e-6, grid, geom, method="xdoppler", iq_storage="float16")


@pytest.mark.parametrize("grid_shape", [(16, 32, 16), (7, 9, 6)])
def test_fp16_opw_on_quantized_point_scene(grid_shape):
    """FP16 channels equal the float64 oracle on the FP16-rounded channels."""
    acq = scene.point_scene(grid_shape=grid_shape)
    q16 = [(x.real.astype(np.float16).astype(np.float32) + 1j * x.imag.astype(np.float16).astype(np.float32)).astype(
        np.complex64) for x in (acq.rc, acq.cr)]
    rounded = scene.Acquisition(q16[0], q16[1], acq.x_el, acq.y_el, acq.angles, acq.t_start, acq.grid, acq.c, acq.fs,
                                acq.fd, acq.fnumber)
    (ra, _, _), (qa, _, _) = _oracle(rounded)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    geom = _geom(acq)
    iq = rb.beamform_ensemble(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom, iq_storage="float16")
    assert rel_l2(iq.reshape(-1, 16), expected) < 1e-5
    pd = rb.ensemble_pd_from_channels(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom, iq_storage="float16")
    assert rel_l2(pd.ravel(), combiners.power_doppler(expected)) < 1e-5
    # Against the float64 oracle of the unrounded channels: the FP16 storage error alone.
    (ra, _, _), (qa, _, _) = _oracle(acq)
    exact = ra.sum(axis=1) + qa.sum(axis=1)
    assert rel_l2(pd.ravel(), combiners.power_doppler(exact)) < 1e-3


def _device_run(acq, storage, coefficients, order=None):
    """rcabeam_opw_mixed_device through the C ABI; returns (IQ (nv, nt) complex64, power (nv,))."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.ensemble import _Geometry, pack_channels

    lib = _lib.load()
    g = _Geometry(acq.angles, acq.t_start, acq.grid, _geom(acq), order=order)
    ns, nc, na, nt = acq.rc.shape
    packed = [pack_channels(dev.to_device(x, torch.complex64), storage) for x in (acq.rc, acq.cr)]
    pd = dev.empty((g.nv, 2), torch.float32)
    signal = dev.empty((g.nv, nt), torch.complex64)
    c, fs, fd, fn = g.constants
    _lib.check(lib.rcabeam_opw_mixed_device(
        dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.scan), dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
        dev.ptr(g.starts), dev.ptr(pd), dev.ptr(signal), _lib.STORAGE[storage], _lib.STORAGE["complex64"],
        coefficients, ns, nc, na, nt, g.nv, c, fs, fd, fn), "opw_mixed")
    dev.synchronize()
    return signal.cpu().numpy(), pd.cpu().numpy()[:, 0]


@pytest.mark.parametrize("fnumber", [None, 1.0])
def test_scattered_voxel_order_stages_windows_in_chunks(fnumber):
    """Voxels in random order: a 64-voxel tile spans the whole volume, so its sample window is
    longer than one LDS chunk and is staged in overlapping pieces (some without any tap)."""
    from rcabeam_amd import _lib

    acq = scene.point_scene(grid_shape=(12, 40, 12))
    acq = scene.Acquisition(acq.rc, acq.cr, acq.x_el, acq.y_el, acq.angles, acq.t_start, acq.grid, acq.c, acq.fs,
                            acq.fd, fnumber)
    nv = 12 * 40 * 12
    order = np.random.default_rng(5).permutation(nv)
    print(f"case: scattered order fnumber={fnumber}", flush=True)
    scan = das.scan_coords(acq.grid)[order]
    (ra, _, _), (qa, _, _) = das.delay_apertures(acq.rc, acq.cr, scan, acq.x_el, acq.y_el, acq.angles, acq.t_start,
                                                 acq.c, acq.fs, acq.fd, fnumber)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    iq, power = _device_run(acq, "float32", _lib.COEF_FLOAT32, order=order)
    assert rel_l2(iq, expected) < TOL
    assert rel_l2(power, combiners.power_doppler(expected)) < TOL


@pytest.mark.parametrize("frames", [16, 200])
@pytest.mark.parametrize("fnumber", [None, 1.0])
def test_single_angle_stages_pack_channel_windows(frames, fnumber):
    """One angle per launch: a stage packs the windows of up to 4 channels.  Coarse 8 x 8 tiles
    (0.85 mm voxels) give windows longer than a stage at fnumber None (staged in chunks), runs of
    live channels cross receive-term fills and the RC -> CR boundary, and 200 frames take the
    smallest stage (22 rows)."""
    acq = scene.point_scene(grid_shape=(16, 8, 16), frames=frames)
    one = scene.Acquisition(acq.rc[:, :, 2:3], acq.cr[:, :, 2:3], acq.x_el, acq.y_el, acq.angles[2:3], acq.t_start,
                            acq.grid, acq.c, acq.fs, acq.fd, fnumber)
    print(f"case: single angle frames={frames} fnumber={fnumber}", flush=True)
    (ra, _, _), (qa, _, _) = _oracle(one)
    args = (one.rc, one.cr, one.angles, one.t_start, one.grid, _geom(one))
    r, q = rb.beamform_apertures(*args)
    assert rel_l2(r.reshape(-1, frames), ra[:, 0]) < TOL
    assert rel_l2(q.reshape(-1, frames), qa[:, 0]) < TOL
    expected = ra[:, 0] + qa[:, 0]
    iq = rb.beamform_ensemble(*args)
    assert rel_l2(iq.reshape(-1, frames), expected) < TOL
    pd = rb.ensemble_pd_from_channels(*args)
    assert rel_l2(pd.ravel(), combiners.power_doppler(expected)) < TOL


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
def test_sixteen_bit_coefficients_error(storage):
    """RCABEAM_COEF_16BIT: one 16-bit coefficient per sample.  Against the float64 oracle of the
    rounded channels the only error is the coefficient rounding (2^-11 FP16, 2^-8 BF16 relative
    per tap, averaged over taps); the float32-accurate form stays at float32 rounding."""
    import torch

    from rcabeam_amd import _lib

    acq = scene.point_scene(grid_shape=(16, 32, 16))
    t = [torch.from_numpy(np.ascontiguousarray(x.view(np.float32))) for x in (acq.rc, acq.cr)]
    dtype = torch.float16 if storage == "float16" else torch.bfloat16
    rounded = [x.to(dtype).to(torch.float32).numpy().view(np.complex64) for x in t]
    acq_r = scene.Acquisition(rounded[0], rounded[1], acq.x_el, acq.y_el, acq.angles, acq.t_start, acq.grid, acq.c,
                              acq.fs, acq.fd, acq.fnumber)
    (ra, _, _), (qa, _, _) = _oracle(acq_r)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    accurate, p_acc = _device_run(acq, storage, _lib.COEF_FLOAT32)
    fast, p_fast = _device_run(acq, storage, _lib.COEF_16BIT)
    assert rel_l2(accurate, expected) < TOL
    e_iq = rel_l2(fast, expected)
    e_pd = rel_l2(p_fast, combiners.power_doppler(expected))
    print(f"{storage}: 16-bit coefficients IQ rel L2 {e_iq:.2e}, power {e_pd:.2e}", flush=True)
    assert e_iq < (5e-4 if storage == "float16" else 4e-3)
    assert e_pd < (1e-3 if storage == "float16" else 8e-3)


@pytest.mark.parametrize("shape", [(13, 20, 11), (9, 7, 17)])
@pytest.mark.parametrize("storage,coefficients", [("float32", 0), ("float16", 0), ("float16", 1), ("bfloat16", 1)])
def test_grid_tiles_equal_voxel_list(storage, coefficients, shape):
    """rcabeam_opw_grid_device against the voxel-list entry point on the same packed channels.  complex64
    channels: the grid kernel (das_grid.hip) accumulates taps in the voxel-list kernel's order, with its
    phasor products rounded independently (rel L2 < 1e-6).  16-bit channels: the grid runs the MFMA
    kernel, whose products are the same exact 16-bit products summed in another order; float32 rounding
    of ~770-term sums in two orders differs by ~2^-24 sqrt(770) ~ 2e-6 relative (bound 5e-6; both stay
    within 1e-5 of the float64 oracle, test_fp16_opw_on_quantized_point_scene).  Grid sizes that are not
    multiples of the patch leave partial tiles, and tile counts that are not multiples of the 4 x 4 x 4
    order blocks leave empty tile slots."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.ensemble import _Geometry, pack_channels

    print(f"case: grid vs list {storage} coef={coefficients} shape={shape}", flush=True)
    acq = scene.point_scene(grid_shape=shape)
    lib = _lib.load()
    g = _Geometry(acq.angles, acq.t_start, acq.grid, _geom(acq))
    ns, nc, na, nt = acq.rc.shape
    packed = [pack_channels(dev.to_device(x, torch.complex64), storage) for x in (acq.rc, acq.cr)]
    c, fs, fd, fn = g.constants
    code = _lib.STORAGE[storage]
    out = []
    for grid in (False, True):
        pd = torch.full((g.nv, 2), float("nan"), device="cuda")
        sig = torch.full((g.nv, nt), float("nan"), dtype=torch.complex64, device="cuda")
        if grid:
            nx, nz, ny = g.shape
            _lib.check(lib.rcabeam_opw_grid_device(
                dev.ptr(packed[0]), dev.ptr(packed[1]), code, coefficients, dev.ptr(g.axes[0]), dev.ptr(g.axes[1]),
                dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles), dev.ptr(g.starts),
                dev.ptr(pd), dev.ptr(sig), 0, 0, ns, nc, na, nt, c, fs, fd, fn), "grid")
        else:
            _lib.check(lib.rcabeam_opw_mixed_device(
                dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.scan), dev.ptr(g.xe), dev.ptr(g.ye),
                dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(pd), dev.ptr(sig), code, 0, coefficients, ns, nc, na,
                nt, g.nv, c, fs, fd, fn), "list")
        dev.synchronize()
        out.append((sig.cpu().numpy(), pd.cpu().numpy()))
    assert np.isfinite(out[1][0]).all() and np.isfinite(out[1][1]).all()
    bound = 1e-6 if storage == "float32" else 5e-6
    assert rel_l2(out[1][0], out[0][0]) < bound
    assert rel_l2(out[1][1], out[0][1]) < bound
    np.testing.assert_array_equal(out[1][1][:, 1], 0.0)


@pytest.mark.parametrize("frames", [16, 200])
def test_grid_apertures_and_xdoppler_equal_voxel_list(frames):
    """Separate-aperture IQ (method 5) of the grid kernel against rcabeam_ensemble_device method 5 on the
    scan list, and XDoppler formed from it by rcabeam_xdoppler_from_apertures_device against method 1
    (reference ensemble.cu:236-315): the same taps summed in the same order per aperture (rel L2 < 1e-6)."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.ensemble import _Geometry, pack_channels

    print(f"case: grid apertures / XDoppler frames={frames}", flush=True)
    acq = scene.random_acquisition(samples=64, channels=8, frames=frames,
                                   grid=(np.linspace(-3e-3, 3e-3, 13).astype(np.float32),
                                         np.linspace(4e-3, 9e-3, 10).astype(np.float32),
                                         np.linspace(-3e-3, 3e-3, 11).astype(np.float32)))
    lib = _lib.load()
    g = _Geometry(acq.angles, acq.t_start, acq.grid, _geom(acq))
    ns, nc, na, nt = acq.rc.shape
    packed = [pack_channels(dev.to_device(x, torch.complex64), "float32") for x in (acq.rc, acq.cr)]
    c, fs, fd, fn = g.constants
    nx, nz, ny = g.shape

    def planes(on_grid):
        pd = torch.full((g.nv, 2), float("nan"), device="cuda")
        sig = torch.full((2 * g.nv, nt), float("nan"), dtype=torch.complex64, device="cuda")
        if on_grid:
            _lib.check(lib.rcabeam_opw_grid_device(
                dev.ptr(packed[0]), dev.ptr(packed[1]), 0, 0, dev.ptr(g.axes[0]), dev.ptr(g.axes[1]),
                dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles), dev.ptr(g.starts),
                dev.ptr(pd), dev.ptr(sig), 0, 5, ns, nc, na, nt, c, fs, fd, fn), "grid")
        else:
            _lib.check(lib.rcabeam_ensemble_device(
                dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.scan), dev.ptr(g.xe), dev.ptr(g.ye),
                dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(pd), dev.ptr(sig), None, None,
                dev.ptr(g.geometry_workspace), ns, nc, na, nt, g.nv, c, fs, fd, fn, 5, 2), "list")
        return pd, sig

    (pd_g, ap_g), (_, ap_l) = planes(True), planes(False)
    assert np.isfinite(ap_g.cpu().numpy()).all()
    assert rel_l2(ap_g.cpu().numpy(), ap_l.cpu().numpy()) < 1e-6
    np.testing.assert_array_equal(pd_g.cpu().numpy(), 0.0)
    xd_pd = torch.full((g.nv, 2), float("nan"), device="cuda")
    xd_sig = torch.full((g.nv, nt), float("nan"), dtype=torch.complex64, device="cuda")
    _lib.check(lib.rcabeam_xdoppler_from_apertures_device(dev.ptr(ap_g), g.nv, nt, dev.ptr(xd_pd),
                                                          dev.ptr(xd_sig)), "xdoppler")
    ref_pd = torch.full((g.nv, 2), float("nan"), device="cuda")
    ref_sig = torch.full((g.nv, nt), float("nan"), dtype=torch.complex64, device="cuda")
    _lib.check(lib.rcabeam_ensemble_device(
        dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.scan), dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
        dev.ptr(g.starts), dev.ptr(ref_pd), dev.ptr(ref_sig), None, None, dev.ptr(g.geometry_workspace), ns, nc,
        na, nt, g.nv, c, fs, fd, fn, 1, 2), "list xdoppler")
    assert rel_l2(xd_sig.cpu().numpy(), ref_sig.cpu().numpy()) < 1e-6
    assert rel_l2(xd_pd.cpu().numpy(), ref_pd.cpu().numpy()) < 1e-6


def test_grid_entry_rejects_unsupported_shapes():
    from rcabeam_amd import _lib

    lib = _lib.load()
    one = 1
    # 17 angles exceed the grid kernel; methods other than 0 and 5 are not grid methods.
    for na, method in ((17, 0), (3, 1)):
        status = lib.rcabeam_opw_grid_device(one * 16, one * 16, 0, 0, 16, 16, 16, 1, 1, 1, 16, 16, 16, 16, 16, 0,
                                             0, method, 8, 1, na, 4, 1540.0, 20e6, 15e6, 1.0)
        assert status == 2, (na, method)


@pytest.mark.parametrize("frames", [1, 3, 8, 66, 200])
def test_opw_fp16_entry_point_matches_quantized_fp32(frames):
    """rcabeam_opw_fp16_device through the C ABI (reference include/rcabeam.h:222-235,
    ensemble.cu:798-809; cases of reference tests/test_opw_fp16.py:12-44): FP16-packed channels give
    the float32 reconstruction of the FP16-rounded channels, at scales 0, 1, 1e-4 and 1e-6 (the last
    puts components into the FP16 subnormal range)."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.ensemble import _Geometry, pack_channels

    lib = _lib.load()
    rng = np.random.default_rng(111)
    shape = (48, 3, 4, frames)
    rc = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    cr = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    el = np.linspace(-0.001, 0.001, 3)
    geom = rb.RCAGeometry(el, el, fs=2e6, f_demod=1e6)
    angles = np.linspace(-0.1, 0.1, 4)
    grid = (np.linspace(-0.001, 0.001, 3), np.array([0.005, 0.007]), np.linspace(-0.001, 0.003, 5))
    g = _Geometry(angles, 1e-6, grid, geom)
    c, fs, fd, fn = g.constants
    ns, nc, na, nt = shape
    for scale in (0.0, 1.0, 1e-4, 1e-6):
        a, b = rc * scale, cr * scale
        q = [(x.real.astype(np.float16).astype(np.float32) + 1j * x.imag.astype(np.float16).astype(np.float32))
             .astype(np.complex64) for x in (a, b)]
        packed = [pack_channels(dev.to_device(x, torch.complex64), "float16") for x in (a, b)]
        pd = dev.empty((g.nv, 2), torch.float32)
        signal = dev.empty((g.nv, nt), torch.complex64)
        _lib.check(lib.rcabeam_opw_fp16_device(
            dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.scan), dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
            dev.ptr(g.starts), dev.ptr(pd), dev.ptr(signal), dev.ptr(g.geometry_workspace), ns, nc, na, nt, g.nv,
            c, fs, fd, fn), "opw_fp16")
        dev.synchronize()
        expected = rb.beamform_ensemble(q[0], q[1], angles, 1e-6, grid, geom).reshape(g.nv, nt)
        got = signal.cpu().numpy()
        power = pd.cpu().numpy()
        if scale == 0.0:
            assert not got.any() and not power.any()
            continue
        # The reference's own bar (tests/test_opw_fp16.py:35-36) at every scale.
        np.testing.assert_allclose(got, expected, rtol=1e-4, atol=max(scale * 2e-5, 1e-12))
        e_iq = rel_l2(got, expected)
        e_pd = rel_l2(power[:, 0], np.mean(np.abs(expected.astype(np.complex128)) ** 2, axis=-1))
        print(f"frames {frames} scale {scale:g}: IQ rel L2 {e_iq:.2e}, power {e_pd:.2e}", flush=True)
        # Samples in the FP16 normal range (scales 1, 1e-4): float32 rounding only.  At 1e-6 the samples are
        # FP16 subnormals with a few significant bits; the products with the hi + lo coefficients there
        # measured 6.7e-6 relative (the reference allows 1e-4).
        bar = 1e-6 if scale >= 1e-4 else 2e-5
        assert e_iq < bar and e_pd < bar, scale
        assert not power[:, 1].any()
        # The geometry workspace is required (reference contract, ensemble.cu:674-675).
        status = lib.rcabeam_opw_fp16_device(
            dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.scan), dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
            dev.ptr(g.starts), dev.ptr(pd), dev.ptr(signal), None, ns, nc, na, nt, g.nv, c, fs, fd, fn)
        assert status == _lib.ERROR_ARGUMENT


def _split_grid_run(acq, geom, scale=1.0):
    """rcabeam_pack_ensemble_split_device + rcabeam_opw_grid_split_device on a whole grid."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.ensemble import _Geometry

    g = _Geometry(acq.angles, acq.t_start, acq.grid, geom)
    ns, nc, na, nt = acq.rc.shape
    raw = [torch.from_numpy(np.ascontiguousarray(x * np.complex64(scale))).cuda() for x in (acq.rc, acq.cr)]
    stats = dev.empty((3,), torch.int32)
    for i, x in enumerate(raw):
        _lib.call("rcabeam_iq_absmax_device", dev.ptr(x), 0, x.numel(), dev.ptr(stats), i)
    packed = [dev.empty((nc, na, ns, nt), torch.complex64) for _ in range(2)]
    for x, pk in zip(raw, packed):
        _lib.call("rcabeam_pack_ensemble_split_device", dev.ptr(x), dev.ptr(pk), dev.ptr(stats), ns, nc, na, nt)
    pd = dev.empty((g.nv, 2), torch.float32)
    signal = dev.empty((g.nv, nt), torch.complex64)
    nx, nz, ny = g.shape
    c, fs, fd, fn = g.constants
    _lib.call("rcabeam_opw_grid_split_device", dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(stats),
              *[dev.ptr(a) for a in g.axes], nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
              dev.ptr(g.starts), dev.ptr(pd), dev.ptr(signal), 0, ns, nc, na, nt, c, fs, fd, fn)
    dev.synchronize()
    return signal.cpu().numpy(), pd.cpu().numpy()


@pytest.mark.parametrize("grid_shape,fnumber", [((16, 32, 16), 1.0), ((13, 20, 11), None)])
def test_split_fp16_grid_matches_oracle(grid_shape, fnumber):
    """Split FP16 channels (scaled hi + lo pairs) on MFMA: float32-accurate OPW against the float64 oracle at
    rel L2 1e-5 (measured 2.5e-6 to 3.4e-6: the MFMA float32 accumulation, not the 22-bit hi + lo
    representation, whose per-term error is ~8e-8), independent of the input scale (power-of-two scaling),
    and exact zeros for zero input."""
    acq = scene.point_scene(grid_shape=grid_shape, fnumber=fnumber)
    geom = _geom(acq)
    print(f"case: split FP16 grid {grid_shape} fnumber={fnumber}", flush=True)
    (ra, _, _), (qa, _, _) = _oracle(acq)
    expected = ra.sum(axis=1) + qa.sum(axis=1)
    for scale in (1.0, 1e-4, 1e3):
        iq, pd = _split_grid_run(acq, geom, scale)
        assert rel_l2(iq, expected * scale) < TOL, scale
        assert rel_l2(pd[:, 0], combiners.power_doppler(expected) * scale**2) < TOL, scale
        assert not pd[:, 1].any()
    zero = scene.Acquisition(np.zeros_like(acq.rc), np.zeros_like(acq.cr), acq.x_el, acq.y_el, acq.angles,
                             acq.t_start, acq.grid, acq.c, acq.fs, acq.fd, fnumber)
    iq, pd = _split_grid_run(zero, geom)
    assert not iq.any() and not pd.any()
