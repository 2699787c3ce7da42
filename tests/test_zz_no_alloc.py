"""rcabeam_temporal_gram_f64_noalloc_device (the exact Gram with its partial sums in a library-owned device
array) against rcabeam_temporal_gram_f64_device (caller workspace), through the C ABI.

Two calls in a row, the second with a smaller input, must equal the workspace entry point bit for bit (both
row counts fit the default row groups); a third call above that limit regroups its rows and must agree to
float64 summation order (rel 1e-13), and the second call's result must not depend on the first call's
leftovers.  No call may change the device's free memory.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _first_use(lib):
    """The first call on a device looks the array's address up (HIP may load the library's code object then,
    which holds the array); later calls allocate nothing.  Make that call before any free-memory reading."""
    import torch

    from rcabeam_amd import _lib

    x = torch.zeros((16, 4), dtype=torch.complex64, device="cuda")
    real = torch.empty((int(lib.rcabeam_gram_real_elements(4)),), dtype=torch.float64, device="cuda")
    _lib.check(lib.rcabeam_temporal_gram_f64_noalloc_device(x.data_ptr(), 16, 4, real.data_ptr(), 0), "noalloc")
    torch.cuda.synchronize()


def _gram(lib, x, rows, frames, noalloc, real=None):
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    if real is None:
        real = torch.full((int(lib.rcabeam_gram_real_elements(frames)),), float("nan"), dtype=torch.float64,
                          device="cuda")
    if noalloc:
        # Nothing but the launches between the free-memory readings.
        torch.cuda.synchronize()
        free = torch.cuda.mem_get_info()[0]
        _lib.check(lib.rcabeam_temporal_gram_f64_noalloc_device(dev.ptr(x), rows, frames, dev.ptr(real), 0), "noalloc")
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free
    else:
        nbytes = int(lib.rcabeam_gram_f64_workspace_bytes(rows, frames))
        work = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
        _lib.check(lib.rcabeam_temporal_gram_f64_device(dev.ptr(x), rows, frames, dev.ptr(real), dev.ptr(work),
                                                        nbytes, 0), "workspace")
    torch.cuda.synchronize()
    return real


def _upper(real, frames):
    """The entries the Gram defines: the upper 16 x 16 blocks of the 2T x 2T real Gram (ld 32 ceil(2T/32))."""
    ld = 32 * ((2 * frames + 31) // 32)
    nb = (2 * frames + 15) // 16
    c = real.cpu().numpy().reshape(-1, ld)
    mask = np.zeros(c.shape, bool)
    for i in range(nb):
        mask[16 * i : 16 * i + 16, 16 * i : 16 * nb] = True
    return c[mask]


def test_noalloc_gram_matches_workspace_entry_point():
    import torch

    from rcabeam_amd import _lib

    lib = _lib.load()
    _first_use(lib)
    gen = torch.Generator(device="cuda").manual_seed(7)
    cases = [(300_000, 200), (5_000, 37)]
    x = [torch.randn((rows, frames), dtype=torch.complex64, device="cuda", generator=gen) for rows, frames in cases]
    results = []
    for xi, (rows, frames) in zip(x, cases):
        print(f"case: noalloc Gram rows={rows} frames={frames}", flush=True)
        got = _gram(lib, xi, rows, frames, noalloc=True)
        want = _gram(lib, xi, rows, frames, noalloc=False)
        g, w = _upper(got, frames), _upper(want, frames)
        assert np.isfinite(g).all()
        np.testing.assert_array_equal(g, w)
        results.append(g)
    # Against float64 NumPy on the small case (exact products, float64 sums).
    a = x[1].cpu().numpy().view(np.float32).astype(np.float64)
    expected = a.T @ a
    frames = cases[1][1]
    ld = 32 * ((2 * frames + 31) // 32)
    full = np.zeros((ld, ld))
    full[: 2 * frames, : 2 * frames] = expected
    mask_vals = _upper(torch.from_numpy(full.ravel()), frames)
    np.testing.assert_allclose(results[1], mask_vals, rtol=1e-12, atol=1e-12 * np.abs(expected).max())


def test_noalloc_gram_regroups_beyond_default_groups():
    """Above 1,638,400 rows at 200 frames the default groups no longer fit: larger groups, same Gram to
    float64 summation order."""
    import torch

    from rcabeam_amd import _lib

    lib = _lib.load()
    _first_use(lib)
    rows, frames = 1_700_000, 200
    gen = torch.Generator(device="cuda").manual_seed(8)
    x = torch.randn((rows, frames), dtype=torch.complex64, device="cuda", generator=gen)
    print(f"case: noalloc Gram rows={rows} frames={frames} (regrouped)", flush=True)
    got = _upper(_gram(lib, x, rows, frames, noalloc=True), frames)
    want = _upper(_gram(lib, x, rows, frames, noalloc=False), frames)
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want)) <= 1e-13 * scale
    assert lib.rcabeam_temporal_gram_f64_noalloc_device(x.data_ptr(), rows, 641, x.data_ptr(), 0) == _lib.ERROR_ARGUMENT
