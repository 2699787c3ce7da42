"""RF -> IQ demodulation oracle (`oracle/demod.py`) pinned by what the mathematics fixes (CPU only).

The reference has no RF path, so there are no reference vectors; instead:
* a carrier tone A cos(2 pi fd t + theta) demodulates to A H(0) e^{i theta} up to the FIR's gain at the
  image frequency -2 fd (closed form: |IQ - A H(0) e^{i theta}| <= A |H(-2 fd)|);
* a tone at fd + delta demodulates to A H(delta) e^{i (2 pi delta t_m + theta)}, H the FIR's (real, even)
  frequency response, up to A |H(-2 fd - delta)|;
* an echo env(t - tau) cos(2 pi fd (t - tau)) comes out with the phase exp(-2j pi fd tau) of the
  reference simulator's IQ (src/rcabeam/sim.py:188);
* linearity, and a brute-force double loop on a tiny case.
A planted mistake (the mixer's sign flipped) fails the carrier-tone check.
"""

import numpy as np

from oracle import demod
from rcabeam_amd.demod import lowpass_taps

FS, FD, D = 60e6, 15e6, 3


def _response(h, f, fs=FS):
    k = np.arange(len(h)) - (len(h) - 1) / 2
    return float(np.sum(h * np.cos(2 * np.pi * f * k / fs)))


def _interior(iq, ntaps):
    margin = (ntaps + D - 1) // D + 1
    return iq[margin:-margin]


def test_carrier_tone_demodulates_to_a_constant():
    h = lowpass_taps(63, FS / (2 * D), FS).astype(np.float64)
    n = np.arange(600)
    t0, amp, theta = 1.3e-6, 0.7, 0.4
    rf = amp * np.cos(2 * np.pi * FD * (t0 + n / FS) + theta)
    iq = _interior(demod.rf_to_iq(rf, h, D, FS, FD, t0), 63)
    bound = amp * abs(_response(h, -2 * FD)) + 1e-12
    assert np.max(np.abs(iq - amp * _response(h, 0.0) * np.exp(1j * theta))) <= bound
    assert abs(_response(h, 0.0) - 1.0) < 1e-6  # unit DC gain (float32 taps)


def test_offset_tone_rotates_at_the_offset():
    h = lowpass_taps(63, FS / (2 * D), FS).astype(np.float64)
    n = np.arange(900)
    t0, amp, theta, delta = 0.0, 1.2, -1.1, 1.7e6
    rf = amp * np.cos(2 * np.pi * (FD + delta) * (t0 + n / FS) + theta)
    iq = demod.rf_to_iq(rf, h, D, FS, FD, t0)
    m = np.arange(len(iq))
    expected = amp * _response(h, delta) * np.exp(1j * (2 * np.pi * delta * (t0 + m * D / FS) + theta))
    sl = slice(22, -22)
    assert np.max(np.abs(iq[sl] - expected[sl])) <= amp * abs(_response(h, -2 * FD - delta)) + 1e-12
    assert _response(h, delta) > 0.99  # inside the passband
    assert abs(_response(h, 25e6)) < 5e-3  # stopband


def test_echo_phase_matches_the_simulator_convention():
    """RF echo env(t - tau) cos(2 pi fd (t - tau)) -> env(t - tau) e^{-2j pi fd tau} (sim.py:188)."""
    h = lowpass_taps(63, FS / (2 * D), FS).astype(np.float64)
    n = np.arange(1200)
    t = n / FS
    tau, sigma = 9.37e-6, 0.6e-6
    env = np.exp(-(((t - tau) / sigma) ** 2))
    rf = env * np.cos(2 * np.pi * FD * (t - tau))
    iq = demod.rf_to_iq(rf, h, D, FS, FD)
    tm = np.arange(len(iq)) * D / FS
    expected = np.exp(-(((tm - tau) / sigma) ** 2)) * np.exp(-2j * np.pi * FD * tau)
    peak = np.argmax(np.abs(expected))
    # Near the peak the 10 MHz low-pass barely touches the ~1 MHz envelope band.
    sl = slice(peak - 10, peak + 10)
    assert np.max(np.abs(iq[sl] - expected[sl])) < 2e-3


def test_linearity_and_brute_force():
    rng = np.random.default_rng(5)
    h = rng.normal(size=7)
    a = rng.normal(size=(40, 2, 3))
    b = rng.normal(size=(40, 2, 3))
    t0 = 0.4e-6
    lhs = demod.rf_to_iq(2 * a - 3 * b, h, 3, FS, FD, t0)
    rhs = 2 * demod.rf_to_iq(a, h, 3, FS, FD, t0) - 3 * demod.rf_to_iq(b, h, 3, FS, FD, t0)
    np.testing.assert_allclose(lhs, rhs, atol=1e-12)
    got = demod.rf_to_iq(a, h, 3, FS, FD, t0)
    for m in range(len(got)):
        for idx in np.ndindex(2, 3):
            total = 0j
            for k in range(7):
                nn = 3 * m + k - 3
                if 0 <= nn < 40:
                    total += 2 * h[k] * a[(nn, *idx)] * np.exp(-2j * np.pi * FD * (t0 + nn / FS))
            assert abs(got[(m, *idx)] - total) < 1e-12


def test_planted_sign_mistake_fails_the_tone_check():
    h = lowpass_taps(63, FS / (2 * D), FS).astype(np.float64)
    n = np.arange(600)
    rf = np.cos(2 * np.pi * FD * n / FS + 0.4)
    wrong = _interior(demod.rf_to_iq(rf, h, D, FS, -FD), 63)  # mixer exp(+2j pi fd t)
    assert np.max(np.abs(wrong - np.exp(0.4j))) > 0.5
