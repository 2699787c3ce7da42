"""Float64 restatement of the reference NumPy delay-and-sum.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`).

Follows `src/rcabeam/sim.py:192-199` (`_scan_coords`) and
`src/rcabeam/sim.py:344-393` (`_delay_numpy`), with two deliberate differences
that make it the exact float64 answer to the problem the C ABI receives:

* every quantity the reference C ABI takes as `float` is rounded to float32
  first (`abi_round`), and
* all arithmetic is float64 / complex128 (the reference rounds the
  interpolation weight to float32 and works in complex64, sim.py:383-386).

Semantics kept from the reference:
* receive delay sqrt(z^2 + (v - e)^2) / c, transmit delay (z cos(a) + u sin(a)) / c,
  with (v, u) = (x, y) for RC and (y, x) for CR (sim.py:369-378);
* sample s = (tau - t0[a]) * fs, valid iff 0 <= s <= ns - 1 (closed upper
  bound), otherwise the tap contributes exactly zero (sim.py:379-381);
* linear interpolation between floor(s) and min(floor(s) + 1, ns - 1)
  (sim.py:382-384);
* rotation by exp(+2j pi fd tau) and the dynamic-aperture mask
  |v - e| <= z / (2 F), F = None or F <= 0 meaning full aperture (sim.py:373, 385).

Pinned in `tests/test_oracle.py` by: golden vectors written by the reference
`_delay_numpy` (`tests/golden/delay_*.npz`), the closed receive-window known
answer [1, 2.5, 3, 0] (reference tests/test_iq_sampling.py:42-53), point-target
focus (reference tests/test_opw.py:15-40), batched-equals-per-frame
(reference tests/test_iq_sampling.py:83-97), linearity, and RC/CR symmetry
under the exchange x <-> y.
"""

from __future__ import annotations

import numpy as np

# Voxel x channel x frame elements per temporary block (bounds oracle memory).
_BLOCK_ELEMENTS = 1 << 19


def abi_round(value) -> np.ndarray:
    """Round to float32 (what crosses the C ABI) and widen back to float64."""
    return np.asarray(value, dtype=np.float32).astype(np.float64)


def scan_coords(grid) -> np.ndarray:
    """Flattened float32 `(x, z, y)` scan coordinates, y fastest (sim.py:192-199)."""
    x, z, y = (np.asarray(axis, dtype=np.float64) for axis in grid)
    coords = np.empty((len(x), len(z), len(y), 3), dtype=np.float32)
    coords[..., 0] = x[:, None, None]
    coords[..., 1] = z[:, None]
    coords[..., 2] = y
    return coords.reshape(-1, 3)


def _full_aperture(fnumber) -> bool:
    return fnumber is None or float(np.float32(fnumber)) <= 0


def delay_volumes(
    iq: np.ndarray,
    scan: np.ndarray,
    elements: np.ndarray,
    angles: np.ndarray,
    starts,
    c: float,
    fs: float,
    fd: float,
    fnumber,
    config: str,
    *,
    channels_out: bool = False,
    aperture_depth=None,
    float32_delays: bool = False,
):
    """Delay one aperture of channel IQ onto voxels (sim.py:344-393).

    Parameters
    ----------
    iq
        Channel IQ `(ns, nc, na)` or `(ns, nc, na, nt)`.
    scan
        `(nv, 3)` scan coordinates `(x, z, y)` in meters.
    elements, angles, starts
        Receive element positions, steering angles and per-angle (or scalar)
        first-sample times.
    c, fs, fd, fnumber
        Sound speed, sampling rate, demodulation frequency and f-number
        (`None` or <= 0: full aperture).
    config
        "RC" (row transmit / column receive) or "CR".
    aperture_depth
        Depth that sizes the receive aperture of every voxel (sim.py:372-373,
        used by two_stage.py:163-165); None: each voxel's own depth.
    float32_delays
        Form receive, transmit and sample positions in float32 with the
        reference's operation order.  The reference's NumPy delay runs in
        float32 whenever its grid axes are float32 arrays, as its two-stage
        reference planes are (two_stage.py:37, 163-165); this reproduces that
        rounding.  Default: float64.

    Returns
    -------
    per_angle
        `(nv, na, nt)` complex128 channel sums per angle.
    per_channel
        `(nv, nc, nt)` complex128 angle sums per channel, or None.
    counts
        `(nv,)` active receive channels per voxel, or None.
    """
    if config not in ("RC", "CR"):
        raise ValueError("config must be 'RC' or 'CR'")
    iq = np.asarray(iq)
    if iq.ndim == 3:
        iq = iq[..., None]
    ns, nc, na, nt = iq.shape
    el = abi_round(elements)
    th = abi_round(angles)
    t0 = np.broadcast_to(abi_round(starts), (na,))
    if ns < 2 or nc < 1 or len(el) != nc or len(th) != na:
        raise ValueError("Expected at least two samples and matching channel/angle dimensions")
    c, fs, fd = (float(np.float32(v)) for v in (c, fs, fd))
    full = _full_aperture(fnumber)
    fn = None if full else float(np.float32(fnumber))
    data = iq.astype(np.complex128)
    coords = np.asarray(scan, dtype=np.float32).astype(np.float64)
    nv = len(coords)
    x, z, y = coords[:, 0], coords[:, 1], coords[:, 2]
    rx_coord, tx_coord = (x, y) if config == "RC" else (y, x)
    per_angle = np.zeros((nv, na, nt), np.complex128)
    per_channel = np.zeros((nv, nc, nt), np.complex128) if channels_out else None
    counts = np.empty(nv, np.float64) if channels_out else None
    channels = np.arange(nc)
    block = max(1, _BLOCK_ELEMENTS // (nc * nt))
    cos_t, sin_t = np.cos(th), np.sin(th)
    if float32_delays:
        f32 = np.float32
        coords32 = np.asarray(scan, dtype=np.float32)
        el32, th32, t032 = el.astype(f32), th.astype(f32), t0.astype(f32)
        c32, fs32 = f32(c), f32(fs)
    for first in range(0, nv, block):
        sl = slice(first, min(first + block, nv))
        depth = z[sl, None]
        distance = rx_coord[sl, None] - el[None, :]
        if float32_delays:
            x32, z32, y32 = coords32[sl, 0], coords32[sl, 1], coords32[sl, 2]
            rx32, tx32 = (x32, y32) if config == "RC" else (y32, x32)
            depth32 = z32[:, None]
            distance32 = rx32[:, None] - el32[None, :]
            receive32 = np.sqrt(depth32**2 + distance32**2) / c32
        receive = np.sqrt(depth**2 + distance**2) / c
        if full:
            active = np.ones_like(distance, dtype=bool)
        elif aperture_depth is None:
            active = np.abs(distance) <= depth / (2 * fn)
        else:
            # A fixed aperture depth is a discrete float32 decision in the reference
            # (two_stage.py:163-165 with float32 plane grids; two_stage.cu:56).
            radius = np.float32(aperture_depth) / (np.float32(2) * np.float32(fn))
            active = np.abs(distance.astype(np.float32)) <= radius
        if counts is not None:
            counts[sl] = active.sum(axis=1)
        for m in range(na):
            transmit = (z[sl] * cos_t[m] + tx_coord[sl] * sin_t[m]) / c
            tau = transmit[:, None] + receive
            sample = (tau - t0[m]) * fs
            phase = 2 * np.pi * fd * tau
            if float32_delays:
                transmit32 = (depth32[:, 0] * np.cos(th32[m]) + tx32 * np.sin(th32[m])) / c32
                tau32 = transmit32[:, None] + receive32
                tau = tau32.astype(np.float64)
                sample = ((tau32 - t032[m]) * fs32).astype(np.float64)
                # The reference's exp(2j pi fd tau) with float32 fd and tau evaluates its argument in complex64:
                # float32(float32(2 pi) fd) times tau, rounded to float32.
                phase = ((np.float32(2 * np.pi) * np.float32(fd)) * tau32).astype(np.float64)
            valid = (sample >= 0) & (sample <= ns - 1)
            sample = np.clip(sample, 0, ns - 1)
            lower = np.floor(sample).astype(np.intp)
            upper = np.minimum(lower + 1, ns - 1)
            w = (sample - lower)[..., None]
            delayed = (1 - w) * data[lower, channels, m] + w * data[upper, channels, m]
            delayed *= (np.exp(1j * phase) * (active & valid))[..., None]
            per_angle[sl, m] = delayed.sum(axis=1)
            if per_channel is not None:
                per_channel[sl] += delayed
    return per_angle, per_channel, counts


def delay_apertures(rc, cr, scan, x_el, y_el, angles, starts, c, fs, fd, fnumber, *, channels_out=False):
    """Delay both apertures; returns `(rc_result, cr_result)` of `delay_volumes` tuples."""
    args = (angles, starts, c, fs, fd, fnumber)
    return (
        delay_volumes(rc, scan, x_el, *args, "RC", channels_out=channels_out),
        delay_volumes(cr, scan, y_el, *args, "CR", channels_out=channels_out),
    )


def opw_iq(rc, cr, scan, x_el, y_el, angles, starts, c, fs, fd, fnumber) -> np.ndarray:
    """OPW IQ `(nv, nt)`: sum over both apertures, all angles and channels."""
    (ra, _, _), (qa, _, _) = delay_apertures(rc, cr, scan, x_el, y_el, angles, starts, c, fs, fd, fnumber)
    return ra.sum(axis=1) + qa.sum(axis=1)
