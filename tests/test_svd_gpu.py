"""Direct temporal SVD on the GPU (rocSOLVER gesvd) and the NumPy-signature clutter filter.

* `rcabeam_amd.svd_clutter_filter` / `temporal_basis` (the reference's package-level NumPy functions,
  src/rcabeam/clutter.py:47-109, method "svd" by default) against the golden vectors the reference's own
  functions wrote (`tests/golden/clutter.npz`: a 40 x 9 complex128 matrix with a strong rank-one
  component; ranks [1, 6) kept, both methods): complex128 results within 1e-10 relative.
* `rcabeam_amd.clutter` (the reference's gpu_clutter.py:43-109, 276-336) with method "svd": singular values,
  projectors (never raw eigenvectors, whose phases are arbitrary) and U Sigma V^H filtered IQ against the
  float64 oracle on complex64 inputs, a stacked RC/CR tuple and a wide matrix (fewer rows than frames).
"""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import clutter

pytestmark = pytest.mark.gpu

GOLDEN = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "clutter.npz")


def _random(rows, frames, seed, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((rows, frames)) + 1j * rng.standard_normal((rows, frames))
    x += 20 * np.outer(rng.standard_normal(rows), np.exp(1j * np.linspace(0, 2, frames)))
    x += 8 * np.outer(rng.standard_normal(rows), np.exp(-1j * np.linspace(0, 3, frames)))
    return x.astype(dtype)


def _projector(basis, start, stop):
    v = np.asarray(basis)[:, start:stop]
    return v @ v.conj().T


@pytest.mark.parametrize("method", ["svd", "gram"])
def test_numpy_api_matches_reference_golden(method):
    x = GOLDEN["x"]
    values, basis = rb.temporal_basis(x, method=method)
    assert values.shape == (9,) and basis.shape == (9, 9)
    assert rel_l2(values, GOLDEN[f"{method}_values"]) < 1e-10
    np.testing.assert_allclose(basis.conj().T @ basis, np.eye(9), atol=1e-12)
    filtered = rb.svd_clutter_filter(x, rank_start=1, rank_stop=6, method=method)
    assert filtered.dtype == np.complex128 and filtered.shape == x.shape
    assert rel_l2(filtered, GOLDEN[f"{method}_filtered"]) < 1e-10
    # complex64 input: complex128 working precision, complex64 output.
    x64 = x.astype(np.complex64)
    f64 = rb.svd_clutter_filter(x64, rank_start=1, rank_stop=6, method=method)
    assert f64.dtype == np.complex64
    expected = clutter.svd_clutter_filter(x64, rank_start=1, rank_stop=6, method="svd")
    assert rel_l2(f64, expected) < 1e-6


def test_numpy_api_default_is_svd_and_validates():
    x = _random(300, 12, 3)
    np.testing.assert_allclose(rb.svd_clutter_filter(x, rank_start=2),
                               rb.svd_clutter_filter(x, rank_start=2, method="svd"), rtol=0, atol=0)
    with pytest.raises(ValueError, match="complex64 or complex128 NumPy array"):
        rb.temporal_basis(x.real.copy())
    bad = x.copy()
    bad[0, 0] = np.nan
    with pytest.raises(ValueError, match="nonempty and finite"):
        rb.svd_clutter_filter(bad, rank_start=1)
    with pytest.raises(ValueError, match="Method must be"):
        rb.temporal_basis(x, method="qr")
    with pytest.raises(ValueError, match="Require 0 <= rank_start"):
        rb.svd_clutter_filter(x, rank_start=5, rank_stop=3)
    empty = rb.svd_clutter_filter(x, rank_start=4, rank_stop=4)
    assert not empty.any()


@pytest.mark.parametrize("rows,frames", [(3000, 40), (10, 24)])
def test_gpu_direct_svd_matches_oracle(rows, frames):
    import torch

    x = _random(rows, frames, rows)
    print(f"case: direct SVD {rows} x {frames}", flush=True)
    ov, ob = clutter.temporal_basis(x, method="svd")
    values, basis = rb.clutter.temporal_basis(torch.from_numpy(x).cuda(), method="svd")
    values, basis = values.cpu().numpy(), basis.cpu().numpy()
    assert values.dtype == np.float64 and basis.shape == (frames, frames)
    k = min(rows, frames)
    assert rel_l2(values, ov) < 1e-12
    assert not values[k:].any()
    np.testing.assert_allclose(basis.conj().T @ basis, np.eye(frames), atol=1e-12)
    for start, stop in ((0, 2), (2, k)):
        assert np.linalg.norm(_projector(basis, start, stop) - _projector(ob, start, stop)) < 1e-9
    filtered = rb.clutter.svd_clutter_filter(torch.from_numpy(x).cuda(), rank_start=2, method="svd").cpu().numpy()
    assert filtered.dtype == np.complex64
    assert rel_l2(filtered, clutter.svd_clutter_filter(x, rank_start=2, method="svd")) < 1e-6
    # complex64 working precision: float32 singular values.
    v32, _ = rb.clutter.temporal_basis(torch.from_numpy(x).cuda(), method="svd", precision="complex64")
    assert v32.dtype == torch.float32 and rel_l2(v32.cpu().numpy(), ov) < 1e-5


def test_gpu_svd_reference_on_rc_cr_tuple():
    """One basis for the stacked RC/CR rows (gpu_clutter.py:57-65); each output keeps its array's shape."""
    import torch

    rc, cr = _random(500, 30, 1).reshape(10, 50, 30), _random(400, 30, 2).reshape(20, 20, 30)
    timings = {}
    (f_rc, f_cr), values, basis = rb.clutter.svd_reference(
        (torch.from_numpy(rc).cuda(), torch.from_numpy(cr).cuda()), rank_start=2, rank_stop=25, timings=timings)
    assert f_rc.shape == rc.shape and f_cr.shape == cr.shape
    assert set(timings) == {"matrix_seconds", "decomposition_seconds", "application_seconds"}
    rows = np.concatenate([rc.reshape(-1, 30), cr.reshape(-1, 30)])
    ov, _ = clutter.temporal_basis(rows, method="svd")
    assert rel_l2(values.cpu().numpy(), ov) < 1e-12
    kept = clutter.svd_clutter_filter(rows, rank_start=2, rank_stop=25, method="svd")
    assert rel_l2(f_rc.cpu().numpy().reshape(-1, 30), kept[:500]) < 1e-6
    assert rel_l2(f_cr.cpu().numpy().reshape(-1, 30), kept[500:]) < 1e-6


def test_numpy_gram_of_complex128_input_uses_complex128_gram():
    """complex128 input on the "gram" method: rocBLAS zherk Gram, float64 EVD, projector filter."""
    x = _random(2000, 16, 9, np.complex128)
    ov, ob = clutter.temporal_basis(x, method="gram")
    values, basis = rb.temporal_basis(x, method="gram")
    assert rel_l2(values, ov) < 1e-10
    assert np.linalg.norm(_projector(basis, 0, 3) - _projector(ob, 0, 3)) < 1e-9
    filtered = rb.svd_clutter_filter(x, rank_start=3, method="gram")
    assert filtered.dtype == np.complex128
    assert rel_l2(filtered, clutter.svd_clutter_filter(x, rank_start=3, method="gram")) < 1e-10
