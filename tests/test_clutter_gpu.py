"""GPU temporal clutter filter against the complex128 oracle (reference tests/test_gpu_clutter.py shape)."""

import numpy as np
import pytest
from conftest import rel_l2

from oracle import clutter as oc

pytestmark = pytest.mark.gpu


def _gpu():
    from rcabeam_amd import clutter

    return clutter


def _spectrum(values, rows=13, seed=47):
    rng = np.random.default_rng(seed)
    frames = len(values)
    left, _ = np.linalg.qr(rng.normal(size=(rows, frames)) + 1j * rng.normal(size=(rows, frames)))
    right, _ = np.linalg.qr(rng.normal(size=(frames, frames)) + 1j * rng.normal(size=(frames, frames)))
    return (left * values) @ right.conj().T, left, right


@pytest.mark.parametrize("shape", [(7,), (3, 7), (19, 6), (2, 4, 6), (2, 1), (16385, 4), (3000, 200), (2000, 224),
                                   (700, 300)])
def test_basis_and_filter_match_oracle(shape):
    """complex128 precision: exact float64 MFMA Gram + float64 EVD (32-row steps up to 224 frames, 16-row
    steps beyond)."""
    gc = _gpu()
    rng = np.random.default_rng(8)
    host = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    original = host.copy()
    timings, diagnostics = {}, {}
    values, basis = gc.temporal_basis(host, timings=timings, diagnostics=diagnostics)
    assert timings["matrix_seconds"] >= 0 and timings["decomposition_seconds"] >= 0
    expected_values, expected_basis = oc.temporal_basis(host.astype(np.complex128), method="svd")
    # Gram + EVD resolves zero singular values only to sqrt(eps * lambda_max) ~ 1e-8 sigma_max
    # (the reference's own GPU test allows atol 1e-7, reference tests/test_gpu_clutter.py:28).
    np.testing.assert_allclose(values, expected_values, rtol=1e-10, atol=1e-7 * expected_values[0])
    frames = shape[-1]
    np.testing.assert_allclose(basis.conj().T @ basis, np.eye(frames), atol=1e-12)
    end = min(3, frames)
    actual = gc.apply_temporal_filter(host, basis, rank_start=0, rank_stop=end)
    expected = oc.svd_clutter_filter(host, rank_start=0, rank_stop=end)
    assert actual.dtype == np.complex64 and actual.shape == host.shape
    np.testing.assert_array_equal(host, original)
    assert rel_l2(actual, expected) < 1e-6
    np.testing.assert_allclose(gc.svd_clutter_filter(host, rank_start=0), host, atol=1e-6)
    np.testing.assert_array_equal(gc.apply_temporal_filter(host, basis, rank_start=frames), np.zeros_like(host))


def test_weak_band_under_strong_clutter():
    """Reference tests/test_clutter.py:79-85: 1e4 : 0.01 dynamic range, complex128 precision."""
    gc = _gpu()
    values = np.array([1e4, 2.0, 1.0, 0.3, 0.1, 0.01])
    iq, left, right = _spectrum(values)
    expected = (left[:, 1:5] * values[1:5]) @ right[:, 1:5].conj().T
    x = iq.astype(np.complex64)
    # complex64 storage of the input itself limits the comparison; compare with the oracle on the same input.
    ref = oc.svd_clutter_filter(x, rank_start=1, rank_stop=5)
    out = gc.svd_clutter_filter(x, rank_start=1, rank_stop=5)
    assert rel_l2(out, ref) < 1e-5
    assert rel_l2(ref, expected) < 1e-2  # the float32 input rounding alone


@pytest.mark.parametrize("start,stop,modes", [(2, 200, 2), (2, 5, 3), (0, 200, 0), (0, 0, 0), (200, 200, 0)])
def test_smaller_projector_and_power(start, stop, modes):
    """Reference tests/test_gpu_clutter.py:95-119 with the identity basis."""
    gc = _gpu()
    rng = np.random.default_rng(12)
    host = (rng.normal(size=(17, 200)) + 1j * rng.normal(size=(17, 200))).astype(np.complex64)
    basis = np.eye(200, dtype=np.complex128)
    expected = np.zeros_like(host)
    expected[:, start:stop] = host[:, start:stop]
    np.testing.assert_allclose(gc.apply_temporal_filter(host, basis, rank_start=start, rank_stop=stop), expected,
                               atol=1e-6)
    power = gc.filtered_power(host, basis, rank_start=start, rank_stop=stop)
    assert power.dtype == np.float64
    np.testing.assert_allclose(power, np.mean(np.abs(expected.astype(np.complex128)) ** 2, axis=-1), rtol=1e-12,
                               atol=1e-14)


def test_residual_power_does_not_subtract_large_energies():
    """Reference tests/test_gpu_clutter.py:160-166."""
    gc = _gpu()
    basis = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]], complex).T / 2
    host = ((1e8 * basis[:, 0] + basis[:, 1])[None, :]).astype(np.complex64)
    np.testing.assert_allclose(gc.filtered_power(host, basis, rank_start=1), oc.filtered_power(host, basis, rank_start=1),
                               rtol=1e-9)


@pytest.mark.parametrize("precision", ["complex64", "complex128"])
@pytest.mark.parametrize("amplitude", [1e38, 1e-40])
def test_scaled_gram_avoids_overflow_and_underflow(precision, amplitude):
    """Reference tests/test_gpu_clutter.py:122-139."""
    gc = _gpu()
    iq = np.full((64, 8), complex(amplitude, amplitude), np.complex64)
    diagnostics = {}
    values, basis = gc.temporal_basis(iq, precision=precision, diagnostics=diagnostics)
    expected, _ = oc.temporal_basis(iq.astype(np.complex128))
    np.testing.assert_allclose(values[0], expected[0], rtol=2e-6)
    assert np.isfinite(values).all() and (values >= 0).all()
    assert diagnostics["input_scale"] == float(iq.real[0, 0])
    power = gc.filtered_power(iq, basis, rank_start=0)
    np.testing.assert_allclose(power, np.full(64, 2 * float(iq.real[0, 0]) ** 2), rtol=1e-6)


def test_split_precision_gram_on_many_rows():
    """complex64 precision: split bfloat16 MFMA Gram; projector of separated modes to 1e-6."""
    gc = _gpu()
    values = np.array([50.0, 20.0, 8.0, 1.0, 0.5, 0.2, 0.1, 0.05])
    iq, left, right = _spectrum(values, rows=20000, seed=3)
    x = iq.astype(np.complex64)
    v_gpu, b_gpu = gc.temporal_basis(x, precision="complex64")
    v_ref, b_ref = oc.temporal_basis(x.astype(np.complex128), method="gram")
    # Each Gram entry carries float32 accumulation error over runs of 2048 rows: an absolute error of
    # ~1e-7 lambda_max^2 in every eigenvalue lambda^2 (as for the reference's complex64 Gram), so weak
    # modes lose relative accuracy while the Gram-level bound holds for all of them.
    np.testing.assert_allclose(v_gpu[:3], v_ref[:3], rtol=1e-5)
    np.testing.assert_allclose(v_gpu ** 2, v_ref ** 2, rtol=0, atol=1e-6 * v_ref[0] ** 2)
    # Projector onto the 3 strong modes: error ~ Gram error / eigen-gap (64 - 1) ~ 1e-5.
    p_gpu = b_gpu[:, :3].astype(np.complex128) @ b_gpu[:, :3].astype(np.complex128).conj().T
    p_ref = b_ref[:, :3] @ b_ref[:, :3].conj().T
    assert np.abs(p_gpu - p_ref).max() < 1e-4


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
def test_sixteen_bit_gram_is_exact(storage):
    """16-bit IQ: one MFMA product is exact, so the Gram equals the float64 Gram of the rounded data."""
    import torch

    gc = _gpu()
    rng = np.random.default_rng(21)
    x = (rng.normal(size=(5000, 24)) + 1j * rng.normal(size=(5000, 24))).astype(np.complex64)
    t = torch.from_numpy(x.view(np.float32).copy()).cuda()
    packed = (t.to(torch.float16) if storage == "float16" else t.to(torch.bfloat16)).contiguous()
    rounded = packed.to(torch.float32).cpu().numpy().view(np.complex64)
    words = packed.view(torch.int32).reshape(5000, 24)
    values, basis = gc.temporal_basis(words, storage=storage)
    expected, _ = oc.temporal_basis(rounded.astype(np.complex128), method="gram")
    # Products are exact; runs of <= 2048 rows are summed in float32, then in float64.
    np.testing.assert_allclose(values.cpu().numpy(), expected, rtol=3e-7)
    power = gc.filtered_power(words, basis, rank_start=2, storage=storage).cpu().numpy()
    _, ref_basis = oc.temporal_basis(rounded.astype(np.complex128))
    assert rel_l2(power, oc.filtered_power(rounded, ref_basis, rank_start=2)) < 1e-6


def test_invalid_inputs():
    gc = _gpu()
    bad = np.ones((4, 3), np.complex64)
    bad[0, 0] = np.nan
    with pytest.raises(ValueError, match="finite"):
        gc.temporal_basis(bad)
    with pytest.raises(ValueError, match="frame count"):
        gc.temporal_basis((np.ones((3, 4), np.complex64), np.ones((3, 5), np.complex64)))
    with pytest.raises(ValueError, match="Method"):
        gc.temporal_basis(np.ones((3, 4), np.complex64), method="qr")
    iq = np.ones((3, 4), np.complex64)
    for basis in (np.ones((4, 4), complex), np.eye(3, dtype=complex), np.full((4, 4), complex(np.nan, 0))):
        with pytest.raises(ValueError, match="Basis"):
            gc.apply_temporal_filter(iq, basis, rank_start=0)
    for start, stop in ((True, None), (-1, 4), (0, 5), (3, 2)):
        with pytest.raises(ValueError, match=r"Rank|rank"):
            gc.svd_clutter_filter(iq, rank_start=start, rank_stop=stop)


@pytest.mark.parametrize("frames", [8, 24, 64, 200])
@pytest.mark.parametrize("dynamic_db", [0, 60])
def test_split_precision_gram_frame_counts(frames, dynamic_db):
    """complex64 precision (split bfloat16 Gram) across frame counts, with and without a rank-3 clutter
    component 60 dB above white noise; the leading singular values against the oracle."""
    gc = _gpu()
    rng = np.random.default_rng(frames + dynamic_db)
    rows = 1536
    x = (rng.normal(size=(rows, frames)) + 1j * rng.normal(size=(rows, frames))) * 10 ** (-dynamic_db / 20)
    clutter = (rng.normal(size=(rows, 3)) + 1j * rng.normal(size=(rows, 3))) @ np.exp(
        1j * rng.uniform(0, 2 * np.pi, size=(3, frames)))
    x = (x + clutter).astype(np.complex64)
    values, basis = gc.temporal_basis(x, precision="complex64")
    expected, _ = oc.temporal_basis(x.astype(np.complex128), method="svd")
    assert np.isfinite(values).all()
    np.testing.assert_allclose(values[:3], expected[:3], rtol=1e-5)
    # Gram-level error bound for the rest (see test_split_precision_gram_on_many_rows).
    np.testing.assert_allclose(values ** 2, expected ** 2, rtol=0, atol=1e-6 * expected[0] ** 2)
