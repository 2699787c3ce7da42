"""Pin the CPU oracle: reference golden vectors, known answers, invariants (no GPU)."""

from pathlib import Path

import numpy as np
import pytest
from conftest import rel_l2

from oracle import combiners, das
from oracle import clutter as oclutter
from oracle import scene

GOLDEN = Path(__file__).resolve().parent / "golden"


def _delay_fixture():
    g = np.load(GOLDEN / "delay.npz")
    grid = (g["grid_x"], g["grid_z"], g["grid_y"])
    return g, grid, das.scan_coords(grid)


@pytest.mark.parametrize("config", ["RC", "CR"])
@pytest.mark.parametrize("fn_name,fnumber", [("f3", 3.0), ("full", None)])
def test_delay_matches_reference_golden(config, fn_name, fnumber):
    """Oracle DAS equals the reference _delay_numpy (complex64) to its rounding."""
    g, grid, scan = _delay_fixture()
    el = g["x_el"] if config == "RC" else g["y_el"]
    pa, pc, cnt = das.delay_volumes(
        g["iq"], scan, el, g["angles"], g["starts"], 1540.0, 20e6, 15e6, fnumber, config, channels_out=True
    )
    ref_a = g[f"{config}_{fn_name}_angles"].reshape(pa.shape)
    ref_c = g[f"{config}_{fn_name}_channels"].reshape(pc.shape)
    assert np.count_nonzero(ref_a) > 0.5 * ref_a.size
    assert rel_l2(pa, ref_a) < 3e-7
    assert rel_l2(pc, ref_c) < 3e-7
    np.testing.assert_array_equal(cnt, g[f"{config}_{fn_name}_counts"].ravel())


def test_f_number_fixture_binds():
    g, _, _ = _delay_fixture()
    assert not np.array_equal(g["RC_f3_counts"], g["RC_full_counts"])
    assert not np.array_equal(g["CR_f3_angles"], g["CR_full_angles"])


def test_closed_receive_window_known_answer():
    """Reference tests/test_iq_sampling.py:42-53: [1, 2.5, 3, 0]."""
    zero = np.zeros(1)
    raw = np.arange(1, 4, dtype=np.complex64).reshape(3, 1, 1)
    scan = das.scan_coords((zero, np.array([0, 1.5, 2, 2.25]), zero))
    pa, _, _ = das.delay_volumes(raw, scan, zero, zero, 0.0, 2.0, 1.0, 0.0, None, "RC")
    np.testing.assert_array_equal(pa.ravel(), [1, 2.5, 3, 0])


def test_batched_frames_equal_independent_frames():
    rng = np.random.default_rng(43)
    raw = (rng.normal(size=(16, 3, 2, 5)) + 1j * rng.normal(size=(16, 3, 2, 5))).astype(np.complex64)
    el = np.array([-0.25, 0.0, 0.5])
    scan = das.scan_coords((el, el + 1.0, el[:2]))
    args = (np.array([-0.1, 0.13]), np.array([0.1, 0.2]), 2.0, 3.0, 0.25, 1.0)
    batch = das.delay_volumes(raw, scan, el, *args, "CR", channels_out=True)
    for t in range(5):
        single = das.delay_volumes(raw[..., t], scan, el, *args, "CR", channels_out=True)
        np.testing.assert_allclose(batch[0][..., t], single[0][..., 0], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(batch[1][..., t], single[1][..., 0], rtol=1e-13, atol=1e-13)


def test_rc_cr_symmetry_under_xy_exchange():
    """CR with (x, y) swapped equals RC: the two apertures are mirror images."""
    rng = np.random.default_rng(5)
    raw = (rng.normal(size=(40, 4, 3, 2)) + 1j * rng.normal(size=(40, 4, 3, 2))).astype(np.complex64)
    el = das.abi_round(np.linspace(-1e-3, 1e-3, 4))
    angles = das.abi_round(np.deg2rad([-5, 0, 5]))
    x, z, y = das.abi_round([-2e-4, 3e-4]), das.abi_round([4e-3, 4.4e-3]), das.abi_round([1e-4, -5e-4, 6e-4])
    rc, _, _ = das.delay_volumes(raw, das.scan_coords((x, z, y)), el, angles, 4.5e-6, 1540, 20e6, 15e6, 1.0, "RC")
    cr, _, _ = das.delay_volumes(raw, das.scan_coords((y, z, x)), el, angles, 4.5e-6, 1540, 20e6, 15e6, 1.0, "CR")
    rc = rc.reshape(2, 2, 3, 3, 2)
    cr = cr.reshape(3, 2, 2, 3, 2).transpose(2, 1, 0, 3, 4)
    np.testing.assert_allclose(rc, cr, rtol=1e-12, atol=1e-12)


def test_linearity():
    rng = np.random.default_rng(6)
    a = (rng.normal(size=(30, 3, 2, 2)) + 1j * rng.normal(size=(30, 3, 2, 2))).astype(np.complex64)
    b = (rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape)).astype(np.complex64)
    el = np.array([-3e-4, 0.0, 3e-4])
    scan = das.scan_coords((el, np.array([2.5e-3, 3e-3]), el))
    args = (np.array([-0.05, 0.05]), 2.8e-6, 1540.0, 20e6, 15e6, 1.0, "RC")
    fa = das.delay_volumes(a, scan, el, *args)[0]
    fb = das.delay_volumes(b, scan, el, *args)[0]
    fab = das.delay_volumes(a.astype(np.complex128) * 2 - 3j * b.astype(np.complex128), scan, el, *args)[0]
    np.testing.assert_allclose(fab, 2 * fa - 3j * fb, rtol=1e-12, atol=1e-12)


def test_point_target_focus():
    """Reference tests/test_opw.py:15-40: the OPW power peaks within one voxel of the scatterer."""
    f0 = 6e6
    el = (np.arange(24) - 23 / 2) * 0.2e-3
    angles = np.deg2rad(np.linspace(-4, 4, 5))
    point = (0.0, 12e-3, 0.0)
    rc = scene.simulate_point(el, el, 1540.0, 4 * f0, f0, angles, point, 360, 8e-6, "RC")
    cr = scene.simulate_point(el, el, 1540.0, 4 * f0, f0, angles, point, 360, 8e-6, "CR")
    grid = (np.linspace(-0.6e-3, 0.6e-3, 9), np.linspace(11.7e-3, 12.3e-3, 9), np.linspace(-0.6e-3, 0.6e-3, 9))
    iq = das.opw_iq(rc, cr, das.scan_coords(grid), el, el, angles, 8e-6, 1540.0, 4 * f0, f0, 1.0)
    pd = combiners.power_doppler(iq).reshape(9, 9, 9)
    peak = np.unravel_index(np.argmax(pd), pd.shape)
    assert all(abs(p - 4) <= 1 for p in peak)


@pytest.mark.parametrize("config", ["RC", "CR"])
def test_simulator_matches_reference_golden(config):
    s = np.load(GOLDEN / "simulate.npz")
    point = tuple(s["point"])
    gauss = scene.simulate_point(s["x_el"], s["y_el"], 1540.0, 20e6, 15e6, s["angles"], point, 96, 4e-6, config)
    band = scene.simulate_point(
        s["x_el"], s["y_el"], 1540.0, 20e6, 15e6, s["angles"], point, 96, 4e-6, config, bandwidth_hz=15e6
    )
    np.testing.assert_allclose(gauss, s[f"{config}_gauss"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(band, s[f"{config}_band"], rtol=0, atol=1e-7)
    pulse = scene.sample_bandlimited_pulse(s["pulse_delays"], 128, 4e-6, 20e6, 15e6)
    np.testing.assert_allclose(pulse, s["pulse"], rtol=0, atol=1e-13)


@pytest.mark.parametrize("samples", [31, 32])
def test_bandlimited_pulse_full_spectrum_closed_form(samples):
    """Reference tests/test_iq_sampling.py:73-83."""
    fs, bw = 20e6, 15e6
    delays = np.array([0.0, 7.25 / fs])
    freq = np.fft.fftfreq(samples, 1 / fs)
    spectrum = np.exp(-4 * (6 * np.log(10) / 20) * (freq / bw) ** 2)
    spectrum *= (1 + np.cos(np.pi * np.clip((2 * abs(freq) - bw) / (fs - bw), 0, 1))) / 2
    expected = np.fft.ifft(spectrum[:, None] * np.exp(-2j * np.pi * freq[:, None] * delays), axis=0).real
    expected *= samples / spectrum.sum()
    np.testing.assert_allclose(scene.sample_bandlimited_pulse(delays, samples, 0, fs, bw), expected, atol=1e-14)


def test_xdoppler_known_answer():
    """Reference tests/test_xdoppler.py:35-44: |2 * conj(7)| = 14 with swapped apertures."""
    iq = np.array([[2], [7]], np.complex64)
    assert np.abs(combiners.xdoppler_signal(iq, np.array([1]), np.array([0])).mean(axis=-1)) == pytest.approx(14)


def test_rc_fmas_matches_literal_double_sum():
    rng = np.random.default_rng(3)
    iq = rng.standard_normal((2, 3, 6, 5)) + 1j * rng.standard_normal((2, 3, 6, 5))
    ri, qi = np.arange(3), np.arange(3, 6)
    brute = sum(combiners.signed_sqrt(iq[..., i, :] * iq[..., j, :]) for i in ri for j in qi)
    np.testing.assert_allclose(combiners.rc_fmas(iq, ri, qi), brute, rtol=1e-12)


def test_dmas_factorised_equals_pairwise():
    rng = np.random.default_rng(7)
    s = rng.standard_normal((5, 8)) + 1j * rng.standard_normal((5, 8))
    st = combiners.signed_sqrt(s)
    y = np.abs(st.sum(axis=-1)) ** 2 - np.sum(np.abs(st) ** 2, axis=-1)
    np.testing.assert_allclose(y, combiners.dmas_bruteforce(s), rtol=1e-12)


@pytest.mark.parametrize("scale", [1e-12, 1.0, 1e12])
def test_dmas_two_coherent_channels_known_answer(scale):
    """Reference tests/test_ensemble_targets.py:123-130: DMAS = 2s, CCF = 1/2, ACF = 1 => s."""
    s = np.full((1, 1), scale, np.complex128)
    y, wc, wa = combiners.dmas_ccf_acf_core(s, s, s, s, ccf_norm=2)
    assert y[0] == pytest.approx(2 * scale) and wc[0] == pytest.approx(0.5) and wa[0] == pytest.approx(1)


def test_combiners_match_reference_golden():
    c = np.load(GOLDEN / "combiners.npz")
    y, wc, wa = combiners.dmas_ccf_acf_core(c["s_rc"], c["s_cr"], c["a_rc"], c["a_cr"], ccf_norm=c["norm"])
    np.testing.assert_allclose(y, c["dmas_y"], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(wc, c["dmas_ccf"], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(wa, c["dmas_acf"], rtol=2e-6, atol=1e-7)
    for k in (2, 3, 4):
        w = combiners.st_sw_weights(c["stsw_iq"], np.arange(4), np.arange(4, 8), k=k)
        np.testing.assert_allclose(w, c[f"stsw_k{k}"], rtol=1e-5, atol=1e-6)
        assert w.max() == pytest.approx(1.0) and w.min() >= 0


def _spectrum_matrix(values, rows=13):
    rng = np.random.default_rng(47)
    frames = len(values)
    left, _ = np.linalg.qr(rng.normal(size=(rows, frames)) + 1j * rng.normal(size=(rows, frames)))
    right, _ = np.linalg.qr(rng.normal(size=(frames, frames)) + 1j * rng.normal(size=(frames, frames)))
    return (left * values) @ right.conj().T, left, right


@pytest.mark.parametrize("method", ["svd", "gram"])
def test_clutter_known_spectrum(method):
    """Reference tests/test_clutter.py:17-36."""
    values = np.array([20.0, 8.0, 3.0, 1.0, 0.2, 0.03])
    matrix, left, right = _spectrum_matrix(values, rows=18)
    s, basis = oclutter.temporal_basis(matrix, method=method)
    np.testing.assert_allclose(s, values, rtol=2e-9)
    expected = (left[:, 1:4] * values[1:4]) @ right[:, 1:4].conj().T
    np.testing.assert_allclose(
        oclutter.svd_clutter_filter(matrix, rank_start=1, rank_stop=4, method=method), expected, atol=1e-12
    )
    power = oclutter.filtered_power(matrix, basis, rank_start=1, rank_stop=4)
    np.testing.assert_allclose(power, np.mean(np.abs(expected) ** 2, axis=-1), rtol=1e-10)
    power_sub = oclutter.filtered_power(matrix, basis, rank_start=1)
    kept = (left[:, 1:] * values[1:]) @ right[:, 1:].conj().T
    np.testing.assert_allclose(power_sub, np.mean(np.abs(kept) ** 2, axis=-1), rtol=1e-10)


def test_clutter_retain_all_and_empty():
    rng = np.random.default_rng(12)
    iq = rng.normal(size=(9, 4)) + 1j * rng.normal(size=(9, 4))
    np.testing.assert_allclose(oclutter.svd_clutter_filter(iq, rank_start=0), iq, atol=1e-13)
    np.testing.assert_array_equal(oclutter.svd_clutter_filter(iq, rank_start=4), np.zeros_like(iq))
    with pytest.raises(ValueError):
        oclutter.rank_interval(4, 3, 2)
    with pytest.raises(ValueError):
        oclutter.rank_interval(4, True, None)


def test_clutter_matches_reference_golden():
    c = np.load(GOLDEN / "clutter.npz")
    for method in ("svd", "gram"):
        s, _ = oclutter.temporal_basis(c["x"], method=method)
        np.testing.assert_allclose(s, c[f"{method}_values"], rtol=1e-12)
        out = oclutter.svd_clutter_filter(c["x"], rank_start=1, rank_stop=6, method=method)
        np.testing.assert_allclose(out, c[f"{method}_filtered"], atol=1e-10)


def test_residual_power_without_cancellation():
    """Reference tests/test_gpu_clutter.py:160-166."""
    basis = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]], complex).T / 2
    x = (1e8 * basis[:, 0] + basis[:, 1])[None, :]
    np.testing.assert_allclose(oclutter.filtered_power(x, basis, rank_start=1), [0.25], atol=1e-14)


def test_scaled_gram_matches_plain_gram():
    rng = np.random.default_rng(9)
    x = (rng.normal(size=(20000, 5)) + 1j * rng.normal(size=(20000, 5))) * 1e3
    gram, scale, rows = oclutter.scaled_gram(x)
    np.testing.assert_allclose(gram * scale**2 * rows, x.conj().T @ x, rtol=1e-12)
