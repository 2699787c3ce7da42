"""RF -> IQ demodulation on MI355X (mix, low-pass FIR, decimate) into the IQ the beamformer takes.

The reference starts from demodulated IQ and lists RF-to-IQ processing as open work (ROADMAP.md:58); its
IQ convention is fixed by its simulator and delay-and-sum (src/rcabeam/sim.py:188, 385): an echo delayed
by tau arrives as env(t - tau) exp(-2j pi fd tau).  `rf_to_iq` produces exactly that from real RF
sampled at `fs_rf`: output sample m at t_start + m * decimation / fs_rf, so the IQ keeps the RF start
time and runs at fs_rf / decimation.  `lowpass_taps` designs the FIR (Hamming-windowed sinc with unit
DC gain).  All arithmetic runs in `rcabeam_rf_to_iq_device`.
"""

from __future__ import annotations

import numpy as np

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib


def lowpass_taps(ntaps: int, cutoff_hz: float, fs_rf: float) -> np.ndarray:
    """Odd-length Hamming-windowed sinc low-pass, normalized to unit DC gain (float32)."""
    if ntaps < 1 or ntaps % 2 == 0 or ntaps > 255:
        raise ValueError("ntaps must be odd and at most 255")
    if not 0 < cutoff_hz < fs_rf / 2:
        raise ValueError("Require 0 < cutoff < fs_rf / 2")
    n = np.arange(ntaps) - (ntaps - 1) / 2
    h = np.sinc(2 * cutoff_hz / fs_rf * n) * np.hamming(ntaps) if ntaps > 1 else np.ones(1)
    return (h / h.sum()).astype(np.float32)


def rf_to_iq(rf, fs_rf: float, f_demod: float, decimation: int, *, taps=None, t_start: float = 0.0,
             ns_iq: int | None = None, packed: bool = False):
    """Complex64 IQ `(ns_iq, channels, angles, frames)` (or the packed `(channels, angles, ns_iq, frames)`)
    from real RF `(ns_rf, channels, angles, frames)` float32 or int16.  `taps` default: 63-tap low-pass at
    fs_rf / (2 decimation).  NumPy input gives NumPy output; GPU tensors stay on the GPU."""
    t = dev.torch()
    host = not dev.is_tensor(rf)
    if host:
        arr = np.asarray(rf)
        if arr.dtype not in (np.float32, np.int16):
            raise ValueError("RF must be float32 or int16")
        x = dev.to_device(arr, t.float32 if arr.dtype == np.float32 else t.int16)
    else:
        if rf.dtype not in (t.float32, t.int16) or not rf.is_cuda:
            raise ValueError("RF must be a float32 or int16 GPU tensor")
        x = rf.contiguous()
    if x.ndim != 4 or any(n == 0 for n in x.shape):
        raise ValueError("RF must be a nonempty (samples, channels, angles, frames) array")
    if int(decimation) < 1:
        raise ValueError("decimation must be positive")
    ns_rf, nc, na, nt = x.shape
    if ns_iq is None:
        ns_iq = ns_rf // int(decimation)
    if taps is None:
        taps = lowpass_taps(63, fs_rf / (2 * int(decimation)), fs_rf)
    h = dev.to_device(np.ascontiguousarray(taps, dtype=np.float32), t.float32)
    shape = (nc, na, ns_iq, nt) if packed else (ns_iq, nc, na, nt)
    out = dev.empty(shape, t.complex64)
    _lib.call("rcabeam_rf_to_iq_device", dev.ptr(x), 0 if x.dtype == t.float32 else 1, ns_rf, nc, na, nt,
              dev.ptr(h), h.numel(), int(decimation), float(fs_rf), float(f_demod), float(t_start), dev.ptr(out),
              int(ns_iq), int(packed))
    return dev.host(out) if host else out
