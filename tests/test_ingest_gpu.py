"""Host ingest through pinned staging (rcabeam_upload_host) against the host bytes, and the
beamforming entry points on host arrays large enough to take it."""

import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count", [1, 4097, (64 << 20) // 8, (64 << 20) // 8 + 3, 3 * (64 << 20) // 8 + 1001])
def test_staged_upload_is_exact(count):
    """Sizes below, at and across the 64 MiB chunk boundary arrive bit for bit."""
    import torch

    from rcabeam_amd import _device as dev

    print(f"case: staged upload of {count} complex64 values", flush=True)
    rng = np.random.default_rng(count % 1000)
    host = (rng.standard_normal(count) + 1j * rng.standard_normal(count)).astype(np.complex64)
    out = dev.upload(host, torch.complex64)
    assert out.dtype == torch.complex64 and out.is_cuda and tuple(out.shape) == (count,)
    np.testing.assert_array_equal(out.cpu().numpy(), host)


def test_staged_upload_converts_dtype_and_rejects_bad_arguments():
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    host = np.arange(10, dtype=np.complex128) * (1 + 2j)
    out = dev.upload(host, torch.complex64)
    np.testing.assert_array_equal(out.cpu().numpy(), host.astype(np.complex64))
    lib = _lib.load()
    assert lib.rcabeam_upload_host(None, None, 0, 1) == 0  # nothing to copy
    assert lib.rcabeam_upload_host(None, out.data_ptr(), 8, 1) == 2
    assert lib.rcabeam_upload_host(host.ctypes.data, out.data_ptr(), 8, 0) == 2


def test_host_input_above_staging_threshold_matches_device_input():
    """An ensemble above the staging threshold (50 MB per aperture) as NumPy and as device tensors."""
    import torch

    import rcabeam_amd as rb
    from rcabeam_amd import _device as dev
    from oracle import scene

    acq = scene.random_acquisition(samples=512, channels=64, angles=3, frames=64, seed=5,
                                   grid=(np.linspace(-2e-3, 2e-3, 8).astype(np.float32),
                                         np.linspace(2e-3, 6e-3, 8).astype(np.float32),
                                         np.linspace(-2e-3, 2e-3, 8).astype(np.float32)))
    assert acq.rc.nbytes >= dev.STAGED_UPLOAD_BYTES
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    timing = {}
    from_host = rb.beamform_ensemble(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom, timings=timing)
    assert timing["raw_upload_seconds"] > 0
    rc_d = torch.from_numpy(acq.rc).cuda()
    cr_d = torch.from_numpy(acq.cr).cuda()
    from_device = rb.beamform_ensemble(rc_d, cr_d, acq.angles, acq.t_start, acq.grid, geom).cpu().numpy()
    # Host input: the CR upload overlaps the RC pass (two aperture passes, one extra float32 addition).
    assert rel_l2(from_host, from_device) < 1e-6
    pd_host = rb.ensemble_pd_from_channels(acq.rc, acq.cr, acq.angles, acq.t_start, acq.grid, geom)
    pd_device = rb.ensemble_pd_from_channels(rc_d, cr_d, acq.angles, acq.t_start, acq.grid, geom).cpu().numpy()
    assert rel_l2(pd_host, pd_device) < 1e-6


def test_aperture_passes_equal_one_grid_pass():
    """rcabeam_opw_grid_aperture_device: the RC pass, then the CR pass added to it, against
    rcabeam_opw_grid_device on the same packed channels (rel L2 1e-6: one extra float32 addition per
    sample), with and without an f-number."""
    import torch

    import rcabeam_amd as rb
    from oracle import scene
    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.ensemble import _Geometry, pack_channels

    for fnumber in (1.0, None):
        acq = scene.point_scene(grid_shape=(16, 24, 16), fnumber=fnumber)
        geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=fnumber)
        g = _Geometry(acq.angles, acq.t_start, acq.grid, geom)
        ns, nc, na, nt = acq.rc.shape
        packed = [pack_channels(torch.from_numpy(x).cuda(), "float32") for x in (acq.rc, acq.cr)]
        nx, nz, ny = g.shape
        c, fs, fd, fn = g.constants
        common = [dev.ptr(a) for a in g.axes] + [nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
                                                 dev.ptr(g.starts)]
        pd1, sig1 = dev.empty((g.nv, 2), torch.float32), dev.empty((g.nv, nt), torch.complex64)
        _lib.call("rcabeam_opw_grid_device", dev.ptr(packed[0]), dev.ptr(packed[1]), 0, 0, *common, dev.ptr(pd1),
                  dev.ptr(sig1), 0, 0, ns, nc, na, nt, c, fs, fd, fn)
        pd2, sig2 = dev.empty((g.nv, 2), torch.float32), dev.empty((g.nv, nt), torch.complex64)
        for aperture in (0, 1):
            _lib.call("rcabeam_opw_grid_aperture_device", dev.ptr(packed[0]), dev.ptr(packed[1]), *common,
                      dev.ptr(pd2), dev.ptr(sig2), aperture, aperture, ns, nc, na, nt, c, fs, fd, fn)
        dev.synchronize()
        assert rel_l2(sig2.cpu().numpy(), sig1.cpu().numpy()) < 1e-6
        assert rel_l2(pd2[:, 0].cpu().numpy(), pd1[:, 0].cpu().numpy()) < 1e-6
        # The RC pass alone is the RC half of the separate-aperture output.
        planes = dev.empty((2 * g.nv, nt), torch.complex64)
        _lib.call("rcabeam_opw_grid_device", dev.ptr(packed[0]), dev.ptr(packed[1]), 0, 0, *common, dev.ptr(pd1),
                  dev.ptr(planes), 0, 5, ns, nc, na, nt, c, fs, fd, fn)
        _lib.call("rcabeam_opw_grid_aperture_device", dev.ptr(packed[0]), dev.ptr(packed[1]), *common,
                  dev.ptr(pd2), dev.ptr(sig2), 0, 0, ns, nc, na, nt, c, fs, fd, fn)
        dev.synchronize()
        assert rel_l2(sig2.cpu().numpy(), planes.reshape(g.nv, 2, nt)[:, 0].cpu().numpy()) < 1e-6


def test_staged_upload_waits_for_earlier_null_stream_work():
    """The staged copies run on their own stream; work already enqueued on the null stream that writes the
    destination must finish first, so the buffer ends up holding exactly the host data."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    count = 3 * (64 << 20) // 8 + 17
    rng = np.random.default_rng(3)
    host = (rng.standard_normal(count) + 1j * rng.standard_normal(count)).astype(np.complex64)
    out = torch.empty(count, dtype=torch.complex64, device="cuda")
    junk = torch.full_like(out, 7.0 - 3.0j)
    for _ in range(200):  # ~tens of ms of null-stream writes into the destination
        out.copy_(junk)
    _lib.call("rcabeam_upload_host", host.ctypes.data, out.data_ptr(), host.nbytes, dev.upload_threads())
    np.testing.assert_array_equal(out.cpu().numpy(), host)
