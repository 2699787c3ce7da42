"""Reusable OPW buffers on one MI355X (reference src/rcabeam/gpu.py:16-209).

`OpwWorkspace` keeps the reference's contract on torch ROCm tensors instead of
CuPy arrays: fixed-shape packed channels and outputs allocated once, `pack`
of device IQ, `reconstruct` without allocation, borrowed outputs that the
next call overwrites, a construction-time metadata snapshot, a per-workspace
lock, and rejection of any device or stream other than the construction
device's default stream (every launch goes to the null stream).
"""

from __future__ import annotations

from operator import index
from threading import Lock

import numpy as np
from numpy.typing import NDArray

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd.ensemble import GRID_MAX_ANGLES, GRID_MAX_FRAMES, IQStorage, opw_voxel_order, pack_channels
from rcabeam_amd.sim import RCAGeometry, _scan_coords


class OpwWorkspace:
    """Own fixed-shape packed channels, scratch and device IQ/power outputs.

    Parameters
    ----------
    channel_shape
        Nonempty `(samples, channels, angles, frames)`; samples must be >= 2.
    angles, t_start
        Steering angles and scalar/per-angle sample starts (radians, seconds).
    grid, geom
        Cartesian `(x, z, y)` coordinate vectors and RCA geometry, copied at
        construction.
    iq_storage
        Packed channel storage: "float32", "float16" (reference) or
        "bfloat16" (extension).  16-bit storage is lossy; arithmetic is
        float32-accurate and the outputs are float32/complex64.
    spatial_order
        Traverse 2 x 4 lateral patches and restore canonical IQ/power on the
        GPU (one extra IQ buffer).

    Notes
    -----
    Calls are synchronous and serialized per workspace, on the default stream
    of the construction device.  `reconstruct` returns a borrowed complex64
    `(nx, nz, ny, frames)` tensor, overwritten by the next call.
    """

    def __init__(
        self,
        channel_shape: tuple[int, int, int, int],
        angles: NDArray[np.floating],
        t_start: float | NDArray[np.floating],
        grid: tuple[NDArray[np.floating], NDArray[np.floating], NDArray[np.floating]],
        geom: RCAGeometry,
        *,
        iq_storage: IQStorage = "float32",
        spatial_order: bool = False,
    ) -> None:
        if len(channel_shape) != 4 or any(index(n) < 1 for n in channel_shape) or channel_shape[0] < 2:
            raise ValueError("Expected nonempty (samples, channels, angles, frames), samples >=2")
        if iq_storage not in ("float32", "float16", "bfloat16"):
            raise ValueError("IQ storage must be float32, float16 or bfloat16")
        self._channel_shape = tuple(index(n) for n in channel_shape)
        ns, nc, na, nt = self._channel_shape
        if any(np.ndim(v) != 1 for v in (angles, geom.x_el, geom.y_el)):
            raise ValueError("Geometry must contain coordinate/angle vectors")
        if len(angles) != na or len(geom.x_el) != nc or len(geom.y_el) != nc:
            raise ValueError("Geometry must match channel/angle dimensions")
        if len(grid) != 3 or any(np.ndim(axis) != 1 or not len(axis) for axis in grid):
            raise ValueError("Grid axes must be nonempty coordinate vectors")
        starts = np.broadcast_to(np.asarray(t_start, dtype=np.float32), (na,))
        scan = _scan_coords(grid)
        vectors = [np.asarray(v, dtype=np.float32) for v in (geom.x_el, geom.y_el, angles, starts)]
        fn = -1.0 if geom.fnumber is None else geom.fnumber
        constants = np.asarray([geom.c, geom.fs, geom.f_demod, fn], dtype=np.float32)
        if not all(np.isfinite(v).all() for v in (scan, *vectors, constants)) or min(constants[:2]) <= 0:
            raise ValueError("Geometry must be finite with positive sound speed/sampling frequency")
        if geom.fnumber is not None and fn <= 0:
            raise ValueError("F-number must be positive or None")
        t = dev.torch()
        self._device = dev.device()
        self._check_context()
        _lib.load()
        self._lock = Lock()
        self._ready = False
        self._has_output = False
        self._storage = iq_storage
        self._constants = tuple(float(v) for v in constants)
        self._shape = (len(grid[0]), len(grid[1]), len(grid[2]))
        nv = len(scan)
        self._nv = nv
        self._inverse = None
        if spatial_order:
            order = opw_voxel_order(self._shape)
            scan = scan[order]
            inverse = np.empty_like(order)
            inverse[order] = np.arange(nv)
            self._inverse = dev.to_device(inverse.astype(np.int64), t.int64)
        self._scan = dev.to_device(np.ascontiguousarray(scan, dtype=np.float32), t.float32)
        self._axes = tuple(dev.to_device(np.asarray(axis, dtype=np.float32), t.float32) for axis in grid)
        self._xe, self._ye, self._angles, self._starts = (dev.to_device(np.ascontiguousarray(v), t.float32)
                                                          for v in vectors)
        words = t.complex64 if iq_storage == "float32" else t.int32
        self._rc = dev.empty((nc, na, ns, nt), words)
        self._cr = dev.empty((nc, na, ns, nt), words)
        self._range_error = dev.empty((1,), t.int32)
        self._pd = dev.empty((nv, 2), t.float32)
        self._signal = dev.empty((nv, nt), t.complex64)
        self._canonical_pd = dev.empty((nv, 2), t.float32) if spatial_order else self._pd
        self._iq = (dev.empty((*self._shape, nt), t.complex64) if spatial_order
                    else self._signal.view(*self._shape, nt))
        dev.synchronize()

    def _check_context(self) -> None:
        """Reject a different device or a non-default stream before any work."""
        t = dev.torch()
        if t.cuda.current_device() != self._device.index or not dev.on_default_stream():
            raise ValueError("Workspace requires its construction device and the default CUDA stream")

    @property
    def power(self):
        """Borrowed canonical unfiltered power `(nx, nz, ny)` float32, overwritten by the next reconstruction."""
        with self._lock:
            if not self._has_output:
                raise RuntimeError("Reconstruct valid channels before reading power")
            return self._canonical_pd[:, 0].reshape(self._shape)

    def pack(self, rc, cr) -> None:
        """Synchronously pack device IQ into the owned buffers, invalidating output.

        `rc`, `cr`: contiguous complex64 tensors of `channel_shape` on the
        workspace device, produced on (or ordered before) the default stream.
        A failed pack leaves reconstruction disabled.
        """
        t = dev.torch()
        with self._lock:
            self._ready = False
            self._has_output = False
            self._check_context()
            for data in (rc, cr):
                if (not isinstance(data, t.Tensor) or tuple(data.shape) != self._channel_shape
                        or data.dtype != t.complex64 or data.device.type != "cuda"):
                    raise ValueError("Inputs must be matching complex64 CUDA IQ arrays")
                if not data.is_contiguous() or data.device.index != self._device.index:
                    raise ValueError("Inputs must be contiguous and on the workspace device")
            pack_channels(rc, self._storage, self._rc, self._range_error)
            pack_channels(cr, self._storage, self._cr, self._range_error)
            dev.synchronize()
            self._ready = True

    def reconstruct(self):
        """Reconstruct the packed ensemble without allocation; returns the borrowed canonical IQ."""
        t = dev.torch()
        with self._lock:
            self._check_context()
            if not self._ready:
                raise RuntimeError("Pack valid channels before reconstructing")
            self._has_output = False
            ns, nc, na, nt = self._channel_shape
            c, fs, fd, fn = self._constants
            if self._inverse is None and na <= GRID_MAX_ANGLES and nt <= GRID_MAX_FRAMES:
                nx, nz, ny = self._shape
                _lib.call(
                    "rcabeam_opw_grid_device", dev.ptr(self._rc), dev.ptr(self._cr), _lib.STORAGE[self._storage],
                    _lib.COEF_FLOAT32, dev.ptr(self._axes[0]), dev.ptr(self._axes[1]), dev.ptr(self._axes[2]),
                    nx, nz, ny, dev.ptr(self._xe), dev.ptr(self._ye), dev.ptr(self._angles), dev.ptr(self._starts),
                    dev.ptr(self._pd), dev.ptr(self._signal), _lib.STORAGE["complex64"], 0, ns, nc, na, nt,
                    c, fs, fd, fn,
                )
            else:
                _lib.call(
                    "rcabeam_opw_mixed_device", dev.ptr(self._rc), dev.ptr(self._cr), dev.ptr(self._scan),
                    dev.ptr(self._xe), dev.ptr(self._ye), dev.ptr(self._angles), dev.ptr(self._starts),
                    dev.ptr(self._pd), dev.ptr(self._signal), _lib.STORAGE[self._storage],
                    _lib.STORAGE["complex64"], _lib.COEF_FLOAT32, ns, nc, na, nt, self._nv, c, fs, fd, fn,
                )
            if self._inverse is not None:
                t.index_select(self._signal, 0, self._inverse, out=self._iq.view(self._nv, nt))
                t.index_select(self._pd, 0, self._inverse, out=self._canonical_pd)
            dev.synchronize()
            self._has_output = True
            return self._iq
