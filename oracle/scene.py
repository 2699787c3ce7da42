"""Seeded synthetic RCA acquisitions (CPU) for parity tests and the CPU baseline.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`).

`simulate_point` restates src/rcabeam/sim.py:131-189 (Gaussian envelope
exp(-((t - tau) / sigma)^2) times exp(-2j pi fd tau)) and
`sample_bandlimited_pulse` restates src/rcabeam/sim.py:80-128.  Both are pinned
by golden vectors written by the reference functions
(`tests/golden/simulate.npz`); the band-limited pulse is further pinned by the
full-spectrum closed form of reference tests/test_iq_sampling.py:73-83.

`point_scene` builds the five-scatterer, -30 dB-noise acquisition of the
64 + 64-element / 200 um / 20 MS/s / 15 MHz configuration (16 frames, 3 angles,
256 samples, 32 x 64 x 32 grid).  Noise convention: circular complex Gaussian
with E|n|^2 = 10^(-30/10) relative to a unit-amplitude scatterer.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def sample_bandlimited_pulse(delays, nsamp: int, t_start: float, fs: float, bandwidth_hz: float) -> np.ndarray:
    """Gaussian-spectrum pulse with cosine anti-alias roll-off (sim.py:80-128)."""
    delays = np.asarray(delays, dtype=np.float64)
    if (
        nsamp < 2
        or delays.ndim != 1
        or not np.all(np.isfinite(delays))
        or not np.all(np.isfinite([t_start, fs, bandwidth_hz]))
        or not 0 < bandwidth_hz < fs
    ):
        raise ValueError("Require finite delays, nsamp >= 2, and 0 < bandwidth_hz < IQ sampling rate")
    freq = np.fft.rfftfreq(nsamp, d=1 / fs)
    sigma_t = 2 * np.sqrt(6 * np.log(10) / 20) / (np.pi * bandwidth_hz)
    spectrum = np.exp(-((np.pi * sigma_t * freq) ** 2))
    roll = np.clip((freq - bandwidth_hz / 2) / ((fs - bandwidth_hz) / 2), 0, 1)
    spectrum *= (1 + np.cos(np.pi * roll)) / 2
    shifted = spectrum[:, None] * np.exp(-2j * np.pi * freq[:, None] * (delays[None, :] - t_start))
    total = 2 * spectrum.sum() - spectrum[0] - (spectrum[-1] if nsamp % 2 == 0 else 0)
    return np.fft.irfft(shifted, n=nsamp, axis=0) * (nsamp / total)


def simulate_point(
    x_el, y_el, c, fs, fd, angles, point, nsamp, t_start, config, *, sigma_t=0.15e-6, bandwidth_hz=None
) -> np.ndarray:
    """Complex64 `(nsamp, n_el, n_angles)` channel IQ of one scatterer (sim.py:131-189)."""
    xp, zp, yp = point
    el = np.asarray(x_el if config == "RC" else y_el, dtype=np.float64)
    v_rx = xp if config == "RC" else yp
    u_tx = yp if config == "RC" else xp
    t = t_start + np.arange(nsamp) / fs
    receive = np.sqrt(zp**2 + (v_rx - el) ** 2) / c
    out = np.zeros((nsamp, len(el), len(angles)), dtype=np.complex64)
    for m, theta in enumerate(angles):
        tau = (zp * np.cos(theta) + u_tx * np.sin(theta)) / c + receive
        if bandwidth_hz is None:
            envelope = np.exp(-(((t[:, None] - tau) / sigma_t) ** 2))
        else:
            envelope = sample_bandlimited_pulse(tau, nsamp, t_start, fs, bandwidth_hz)
        out[:, :, m] = envelope * np.exp(-2j * np.pi * fd * tau)[None, :]
    return out


@dataclass(frozen=True)
class Acquisition:
    """One seeded RC/CR ensemble plus the geometry that crosses the C ABI."""

    rc: np.ndarray
    cr: np.ndarray
    x_el: np.ndarray
    y_el: np.ndarray
    angles: np.ndarray
    t_start: float
    grid: tuple
    c: float
    fs: float
    fd: float
    fnumber: float | None


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def point_scene(
    *,
    elements: int = 64,
    samples: int = 256,
    frames: int = 16,
    grid_shape=(32, 64, 32),
    angles_deg=(-5.0, 0.0, 5.0),
    scatterers: int = 5,
    noise_db: float = -30.0,
    seed: int = 42,
    t_start: float = 0.0,
    fnumber: float | None = 1.0,
) -> Acquisition:
    """Five seeded point scatterers plus circular Gaussian noise.

    Positions ~ U(x, y in [-5, 5] mm, z in [6, 14] mm), unit amplitude; each
    frame multiplies the static echo by a seeded unit phasor so the ensemble
    is not rank one.  Geometry values are float32-representable so the oracle
    and the float32 C ABI see identical numbers.
    """
    rng = np.random.default_rng(seed)
    pitch = 200e-6
    el = _f32((np.arange(elements) - (elements - 1) / 2) * pitch)
    c, fs, fd = 1540.0, 20e6, 15e6
    angles = _f32(np.deg2rad(np.asarray(angles_deg)))
    nx, nz, ny = grid_shape
    grid = (_f32(np.linspace(-6.4e-3, 6.4e-3, nx)), _f32(np.linspace(5e-3, 15e-3, nz)), _f32(np.linspace(-6.4e-3, 6.4e-3, ny)))
    points = rng.uniform([-5e-3, 6e-3, -5e-3], [5e-3, 14e-3, 5e-3], (scatterers, 3))
    frame_phase = np.exp(1j * rng.uniform(-np.pi, np.pi, (scatterers, frames)))
    out = []
    for config in ("RC", "CR"):
        static = np.stack(
            [simulate_point(el, el, c, fs, fd, angles, tuple(p), samples, t_start, config) for p in points], axis=-1
        ).astype(np.complex128)
        data = np.einsum("scap,pt->scat", static, frame_phase)
        sigma = np.sqrt(10 ** (noise_db / 10) / 2)
        data = data + sigma * (rng.standard_normal(data.shape) + 1j * rng.standard_normal(data.shape))
        out.append(np.ascontiguousarray(data.astype(np.complex64)))
    return Acquisition(out[0], out[1], el, el.copy(), angles, t_start, grid, c, fs, fd, fnumber)


def random_acquisition(
    *, samples=16, channels=2, angles=4, frames=3, grid=None, seed=44, fnumber=1.0, fs=1e6, fd=0.1e6, c=1540.0
) -> Acquisition:
    """Small random-IQ ensemble in the style of reference tests/test_ensemble.py:31-44."""
    rng = np.random.default_rng(seed)
    shape = (samples, channels, angles, frames)
    rc = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    cr = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    x_el = _f32(np.linspace(-0.001, 0.001, channels)) if channels > 1 else _f32([0.001])
    y_el = _f32(np.linspace(-0.0008, 0.0007, channels)) if channels > 1 else _f32([0.0007])
    th = _f32(np.linspace(-0.08, 0.09, angles))
    if grid is None:
        grid = (_f32([0.0, 0.008]), _f32([0.001, 0.007, 0.03]), _f32([0.0]))
    return Acquisition(rc, cr, x_el, y_el, th, 1e-6, tuple(_f32(g) for g in grid), c, fs, fd, fnumber)
