"""Float64 RF -> IQ demodulation: the oracle of `rcabeam_rf_to_iq_device`.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`).

The reference has no RF path (ROADMAP.md:58 lists RF-to-IQ as open work); the convention restated here is
the one its simulator and delay-and-sum fix (src/rcabeam/sim.py:188, 385):
    iq[m] = sum_k 2 h[k] rf[m D + k - c] exp(-2j pi fd t_{m D + k - c}),  t_n = t0 + n / fs,
with c = (K - 1) / 2 and RF outside [0, N) zero.  Pinned in tests/test_demod.py by what the mathematics
fixes rather than by a reference: a carrier tone demodulates to a constant 2 |H(0)| A / 2 = A, a tone at
fd + delta rotates at delta, linearity, the FIR passband / stopband gains, and a brute-force loop.
"""

from __future__ import annotations

import numpy as np


def rf_to_iq(rf, taps, decimation: int, fs: float, fd: float, t0: float = 0.0, ns_iq: int | None = None):
    """Complex128 IQ `(ns_iq, ...)` of real RF `(ns_rf, ...)` (first axis = samples)."""
    rf = np.asarray(rf, dtype=np.float64)
    h = np.asarray(taps, dtype=np.float64)
    k_count = len(h)
    if k_count % 2 == 0:
        raise ValueError("odd tap count required")
    c = (k_count - 1) // 2
    ns = rf.shape[0]
    if ns_iq is None:
        ns_iq = ns // decimation
    n = np.arange(ns)
    mixed = 2 * rf * np.exp(-2j * np.pi * fd * (t0 + n / fs)).reshape((-1,) + (1,) * (rf.ndim - 1))
    padded = np.concatenate([np.zeros((c,) + rf.shape[1:]), mixed, np.zeros((c,) + rf.shape[1:])])
    out = np.zeros((ns_iq,) + rf.shape[1:], np.complex128)
    for m in range(ns_iq):
        window = padded[m * decimation : m * decimation + k_count]  # rf[m D - c .. m D + c]
        out[m] = np.tensordot(h, window, axes=(0, 0))
    return out
