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
    np.testing.assert_array_equal(from_host, from_device)
    assert rel_l2(from_host, from_device) == 0.0


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
