"""Geometry, scan coordinates and single-frame delay-and-sum on the GPU.

Mirrors the public surface of reference src/rcabeam/sim.py: `RCAGeometry`
(:12-37), `depth_axis` (:40-77), `_scan_coords` (:192-199),
`delay_rca_channels` (:202-265) and `delay_rca_channel_data` (:268-337).
The delay functions always run the HIP kernels behind
`rcabeam_delay_rca_channels_device` / `rcabeam_delay_rca_channel_data_device`;
`use_cuda` is accepted for signature compatibility and must stay True (there
is no CPU path in this package).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np
from numpy.typing import NDArray

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib


@dataclass(frozen=True)
class RCAGeometry:
    """Flat row-column-array geometry (reference src/rcabeam/sim.py:12-37).

    x_el: x positions of the column (RC receive) elements, m.
    y_el: y positions of the row (CR receive) elements, m.
    c: sound speed m/s; fs: IQ sampling rate Hz; f_demod: demodulation Hz;
    fnumber: receive f-number, `None` for the full aperture.
    """

    x_el: np.ndarray
    y_el: np.ndarray
    c: float = 1540.0
    fs: float = 24e6
    f_demod: float = 6e6
    fnumber: float | None = 1.0


def depth_axis(
    start: float, stop: float, frequency: float, *, sound_speed: float = 1540.0, spacing: float | None = None
) -> NDArray[np.float64]:
    """Uniform depths including both ends, spacing <= lambda/2 by default (sim.py:40-77)."""
    if (
        not np.isfinite([start, stop, frequency, sound_speed]).all()
        or start < 0
        or stop < start
        or min(frequency, sound_speed) <= 0
    ):
        raise ValueError("Require finite ordered nonnegative depths and positive frequency/sound speed")
    step = sound_speed / (2 * frequency) if spacing is None else spacing
    if not np.isfinite(step) or step <= 0:
        raise ValueError("Depth spacing must be finite and positive")
    return np.linspace(start, stop, int(np.ceil((stop - start) / step)) + 1)


def _scan_coords(grid) -> np.ndarray:
    """Flattened float32 `(x, z, y)` coordinates, voxel id (ix*nz + iz)*ny + iy (sim.py:192-199)."""
    x, z, y = grid
    coords = np.empty((len(x), len(z), len(y), 3), dtype=np.float32)
    coords[..., 0] = np.asarray(x)[:, None, None]
    coords[..., 1] = np.asarray(z)[:, None]
    coords[..., 2] = np.asarray(y)
    return coords.reshape(-1, 3)


def _fnumber(geom: RCAGeometry) -> float:
    return -1.0 if geom.fnumber is None else float(geom.fnumber)


def _frame_inputs(iq_ch, angles, t_start, grid, geom, config):
    if config not in ("RC", "CR"):
        raise ValueError("config must be 'RC' or 'CR'")
    t = dev.torch()
    arr = dev.to_device(iq_ch, t.complex64)
    if arr.ndim != 3:
        raise ValueError("Channel IQ must be (samples, channels, angles)")
    ns, nc, na = arr.shape
    elements = geom.x_el if config == "RC" else geom.y_el
    if ns < 2 or len(elements) != nc or len(angles) != na:
        raise ValueError("Expected at least two samples and matching channel/angle dimensions")
    starts = np.ascontiguousarray(np.broadcast_to(np.asarray(t_start, dtype=np.float32), (na,)))
    scan = dev.to_device(_scan_coords(grid), t.float32)
    el = dev.to_device(np.asarray(elements, dtype=np.float32), t.float32)
    th = dev.to_device(np.asarray(angles, dtype=np.float32), t.float32)
    st = dev.to_device(starts, t.float32)
    return arr, scan, el, th, st, (ns, nc, na)


def delay_rca_channels(
    iq_ch, angles, t_start, grid, geom: RCAGeometry, config: Literal["RC", "CR"], *, use_cuda: bool = True
) -> np.ndarray:
    """Per-angle DAS volumes `(nx, nz, ny, n_angles)` of one frame (sim.py:202-265)."""
    if not use_cuda:
        raise ValueError("rcabeam_amd has no CPU path; use_cuda must be True")
    t = dev.torch()
    arr, scan, el, th, st, (ns, nc, na) = _frame_inputs(iq_ch, angles, t_start, grid, geom, config)
    nv = scan.shape[0]
    out = dev.empty((nv, na), t.complex64)
    _lib.call(
        "rcabeam_delay_rca_channels_device",
        dev.ptr(arr), dev.ptr(scan), dev.ptr(el), dev.ptr(th), dev.ptr(st), dev.ptr(out),
        ns, nc, na, nv, 0 if config == "RC" else 1,
        float(geom.c), float(geom.fs), float(geom.f_demod), _fnumber(geom),
    )
    shape = tuple(len(g) for g in grid)
    return dev.host(out).reshape((*shape, na))


def delay_rca_channel_data(
    iq_ch, angles, t_start, grid, geom: RCAGeometry, config: Literal["RC", "CR"], *, use_cuda: bool = True
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Channel sums `(.., n_channels)`, angle sums `(.., n_angles)`, active counts (sim.py:268-337)."""
    if not use_cuda:
        raise ValueError("rcabeam_amd has no CPU path; use_cuda must be True")
    t = dev.torch()
    arr, scan, el, th, st, (ns, nc, na) = _frame_inputs(iq_ch, angles, t_start, grid, geom, config)
    nv = scan.shape[0]
    channel_sum = dev.empty((nv, nc), t.complex64)
    angle_sum = dev.empty((nv, na), t.complex64)
    n_active = dev.empty((nv,), t.float32)
    _lib.call(
        "rcabeam_delay_rca_channel_data_device",
        dev.ptr(arr), dev.ptr(scan), dev.ptr(el), dev.ptr(th), dev.ptr(st),
        dev.ptr(channel_sum), dev.ptr(angle_sum), dev.ptr(n_active),
        ns, nc, na, nv, 0 if config == "RC" else 1,
        float(geom.c), float(geom.fs), float(geom.f_demod), _fnumber(geom),
    )
    shape = tuple(len(g) for g in grid)
    return (
        dev.host(channel_sum).reshape((*shape, nc)),
        dev.host(angle_sum).reshape((*shape, na)),
        dev.host(n_active).reshape(shape),
    )
