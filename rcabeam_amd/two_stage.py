"""Two-stage OPW approximation on MI355X: dense reference planes + cubic-spline expansion.

Experimental, as in the reference (src/rcabeam/two_stage.py:1-5): the voxel
mapping offset * tan(angle / 2) is a closest-element approximation, not
conventional OPW.  Each aperture is beamformed onto reference planes (receive
line x uniformly spaced reference depths, one plane per elevation slab), a
not-a-knot cubic spline is solved along depth, and every voxel is expanded from
it.  With an F-number the depth axis is cut into bands inside which no receive
mask changes; planes are rebuilt per band, copying the knots of lines whose
channel set did not change.

`TwoStageOpwWorkspace` and `TwoStageStswWorkspace` keep the reference's
workspace contract (two_stage.py:201-420) on torch tensors: geometry
snapshotted at construction, synchronous `reconstruct` on the construction
device's default stream, borrowed outputs overwritten by the next call, a
per-workspace lock.  The band loop and the four kernels run natively
(`rcabeam_two_stage_reconstruct_device` and friends, include/rcabeam_mi355x.h).
"""

from __future__ import annotations

import ctypes
import warnings
from operator import index
from threading import Lock

import numpy as np

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd.sim import RCAGeometry

# Frames per pass of the St-SW workspace: one wave covers up to 256 frames.
STSW_FRAME_CHUNK = 256


def _plane_layout(axis: np.ndarray, n_planes: int):
    """Slab centres of equal-width elevation slabs, each coordinate's nearest slab
    and its offset from that centre (reference two_stage.py:20-25)."""
    lo, hi = float(axis.min()), float(axis.max())
    step = (hi - lo) / n_planes
    centers = (lo + (np.arange(n_planes) + 0.5) * step).astype(np.float32)
    nearest = np.abs(axis[:, None] - centers[None, :]).argmin(axis=1).astype(np.int32)
    return centers, nearest, (axis - centers[nearest]).astype(np.float32)


class _Plan:
    """Validated geometry snapshot and reference-depth axis (reference two_stage.py:28-80)."""

    def __init__(self, shape, angles, t_start, grid, geom, samples_per_wavelength, n_planes, baseband):
        if isinstance(n_planes, (bool, np.bool_)) or index(n_planes) < 1:
            raise ValueError("Require at least one reference plane")
        if len(shape) != 4 or any(index(n) < 1 for n in shape) or shape[0] < 2:
            raise ValueError("Expected nonempty (samples, channels, angles, frames), samples >=2")
        self.shape = tuple(index(n) for n in shape)
        vectors = (*grid, angles, geom.x_el, geom.y_el)
        if len(grid) != 3 or any(np.ndim(v) != 1 or not len(v) for v in vectors):
            raise ValueError("Grid, angles and elements must be nonempty vectors")
        self.axes = tuple(np.array(v, dtype=np.float32) for v in grid)
        self.angles = np.array(angles, dtype=np.float32)
        self.elements = (np.array(geom.x_el, dtype=np.float32), np.array(geom.y_el, dtype=np.float32))
        if len(self.angles) != self.shape[2] or any(len(e) != self.shape[1] for e in self.elements):
            raise ValueError("Geometry must match channel and angle dimensions")
        self.starts = np.array(np.broadcast_to(np.asarray(t_start, dtype=np.float32), self.angles.shape))
        fn = -1.0 if geom.fnumber is None else geom.fnumber
        spw = float(samples_per_wavelength)
        constants = np.asarray([geom.c, geom.fs, geom.f_demod, fn, spw], dtype=np.float32)
        if not all(np.isfinite(v).all() for v in (*self.axes, self.angles, *self.elements, self.starts, constants)):
            raise ValueError("Geometry and sampling must be finite")
        self.n_planes = index(n_planes)
        if self.n_planes * max(len(self.axes[0]), len(self.axes[2])) > np.iinfo(np.int32).max:
            raise ValueError("Reference plane line count must fit signed int32")
        if min(constants[:3]) <= 0 or spw < 2 or (geom.fnumber is not None and fn <= 0):
            raise ValueError("Require positive c/fs/carrier/F-number and at least two axial samples per wavelength")
        # Broadside two-way delay 2z/c: the highest represented RF frequency is
        # f_demod + fs/2, which needs more than 4 (1 + fs / (2 f_demod)) plane
        # samples per wavelength unless the carrier is removed first.
        density = 4 * (1 + geom.fs / (2 * geom.f_demod))
        if not baseband and spw <= density:
            warnings.warn(
                f"Axial undersampling experiment: use more than {density:g} plane samples/wavelength "
                "to resolve the full sampled IQ band at broadside. The carrier alone requires more than four.",
                UserWarning, stacklevel=3)
        x, z, y = self.axes
        if np.any(z <= 0) or np.any(np.abs(self.angles) >= np.pi / 2):
            raise ValueError("Require positive depths and steering angles strictly within +/-pi/2")
        self.layouts = (_plane_layout(x, self.n_planes), _plane_layout(y, self.n_planes))
        self.slopes = np.tan(self.angles / np.float32(2)).astype(np.float32)
        self.shifts = np.concatenate([(off[:, None] * self.slopes[None, :]).ravel() for _, _, off in self.layouts])
        self.dz = np.float32(geom.c / geom.f_demod / spw)
        if not np.isfinite(self.dz) or self.dz <= 0:
            raise ValueError("Reference-depth spacing is not representable in float32")
        two = np.float32(2)
        z0 = np.float32(z.min() + self.shifts.min() - two * self.dz)
        end = np.float32(z.max() + self.shifts.max() + two * self.dz)
        if not np.isfinite([z0, end]).all():
            raise ValueError("Mapped reference depths must be finite")
        if z0 <= 0:
            raise ValueError("Mapped reference planes reach nonpositive depths; reduce steering/FoV or increase depth")
        npz = int(np.ceil(np.float32(end - z0) / self.dz)) + 1
        if max(*self.shape, *(len(v) for v in self.axes), npz) > np.iinfo(np.int32).max:
            raise ValueError("Prototype dimensions must fit signed int32")
        self.plane_z = (z0 + np.arange(npz, dtype=np.float32) * self.dz).astype(np.float32)
        self.c, self.fs, self.fd, self.fnumber = (np.float32(v) for v in constants[:4])
        self.bands = self._bands()

    def _bands(self):
        """`(first depth, depth count, first knot, knot count, aperture radius)` per run of
        depths sharing every receive mask (reference two_stage.py:83-102)."""
        x, z, y = self.axes
        pz, shifts = self.plane_z, self.shifts
        if self.fnumber <= 0:
            return [(0, len(z), 0, len(pz), np.float32(-1))]
        # Masks are nested in the radius, so equal total counts over both
        # apertures mean equal masks on every line.
        pairs = np.sort(np.concatenate([np.abs(a[:, None] - e[None, :]).ravel()
                                        for a, e in zip((x, y), self.elements)]))
        radius = z / (np.float32(2) * self.fnumber)
        counts = np.searchsorted(pairs, radius, side="right")
        starts = np.concatenate([[0], np.flatnonzero(np.diff(counts)) + 1, [len(z)]])
        out = []
        for first, stop in zip(starts[:-1], starts[1:]):
            depths = z[first:stop]
            lo = max(0, int(np.searchsorted(pz, depths.min() + shifts.min())) - 3)
            hi = min(len(pz), int(np.searchsorted(pz, depths.max() + shifts.max(), side="right")) + 3)
            out.append((int(first), int(stop - first), lo, hi - lo, np.float32(radius[first])))
        return out

    def unchanged(self, axis: np.ndarray, elements: np.ndarray) -> np.ndarray:
        """uint8 `(bands, lines)`: the line's channel count (hence set) equals the previous band's
        (reference two_stage.py:105-116)."""
        radii = np.array([b[4] for b in self.bands], dtype=np.float32)
        flags = np.zeros((len(self.bands), len(axis)), np.uint8)
        if len(self.bands) > 1:
            dist = np.sort(np.abs(axis[:, None] - elements[None, :]), axis=1)
            counts = np.stack([np.searchsorted(row, radii, side="right") for row in dist], axis=1)
            flags[1:] = counts[1:] == counts[:-1]
        return flags


def _spline_factors(n: int):
    """Thomas factors of the rows M(i-1) + 4 M(i) + M(i+1) after M1 is known:
    u[1] = 0, inv[i] = 1 / (4 - u[i-1]), u[i] = inv[i] (reference two_stage.py:119-126)."""
    upper = np.zeros(max(n, 4), np.float64)
    inverse = np.zeros(max(n, 4), np.float64)
    for i in range(2, len(upper)):
        inverse[i] = 1.0 / (4.0 - upper[i - 1])
        upper[i] = inverse[i]
    return upper.astype(np.float32), inverse.astype(np.float32)


class TwoStageOpwWorkspace:
    """Opt-in GPU two-stage OPW: synchronous, borrowed canonical IQ, no input packing.

    Parameters follow the reference (two_stage.py:213-225): `channel_shape`
    `(samples, channels, angles, frames)`, `angles`, `t_start`, `grid` `(x, z, y)`,
    `geom`; `samples_per_wavelength` reference depths per carrier wavelength;
    `frame_chunk` frames per pass (bounds the plane buffers); `n_planes`
    elevation slabs per aperture; `baseband` removes the closest-element carrier
    before the spline and restores it at the voxel.  Inputs are complex64
    device tensors; there is no clutter filter or 16-bit path.
    """

    _stsw = False

    def __init__(self, channel_shape, angles, t_start, grid, geom: RCAGeometry, *, samples_per_wavelength=8,
                 frame_chunk=32, n_planes=1, baseband=False):
        plan = _Plan(channel_shape, angles, t_start, grid, geom, samples_per_wavelength, n_planes, baseband)
        if isinstance(frame_chunk, (bool, np.bool_)) or index(frame_chunk) < 1:
            raise ValueError("Frame chunk must be positive")
        t = dev.torch()
        self._device = dev.device()
        self._check_context()
        _lib.load()
        self._lock = Lock()
        self._plan = plan
        ns, nc, na, nt = plan.shape
        self._shape = plan.shape
        self._chunk = min(STSW_FRAME_CHUNK if self._stsw else index(frame_chunk), nt)
        x, z, y = plan.axes
        nx, nz, ny = len(x), len(z), len(y)
        bands = plan.bands
        if min(b[3] for b in bands) < 4:
            raise ValueError("Every depth band needs at least four reference depths")
        f32 = t.float32

        def put(a, dtype=f32):
            return dev.to_device(np.ascontiguousarray(a), dtype)

        # Host band table (kept alive: the plan struct points at it).
        self._host = {
            "depth_start": np.array([b[0] for b in bands], np.int32),
            "depth_count": np.array([b[1] for b in bands], np.int32),
            "knot_first": np.array([b[2] for b in bands], np.int32),
            "knot_count": np.array([b[3] for b in bands], np.int32),
            "radius": np.array([b[4] for b in bands], np.float32),
            "knots": plan.plane_z.copy(),
        }
        upper, inverse = _spline_factors(int(self._host["knot_count"].max()))
        (xc, xp, xo), (yc, yp, yo) = plan.layouts
        self._dev = {
            "x": put(x), "z": put(z), "y": put(y), "x_el": put(plan.elements[0]), "y_el": put(plan.elements[1]),
            "angles": put(plan.angles), "starts": put(plan.starts), "slopes": put(plan.slopes),
            "knots": put(plan.plane_z), "x_centers": put(xc), "x_plane": put(xp, t.int32), "x_offset": put(xo),
            "y_centers": put(yc), "y_plane": put(yp, t.int32), "y_offset": put(yo),
            "x_same": put(plan.unchanged(x, plan.elements[0]), t.uint8),
            "y_same": put(plan.unchanged(y, plan.elements[1]), t.uint8),
            "upper": put(upper), "inverse": put(inverse),
        }
        if baseband:
            # Closest-element carrier k (1 + cos a) z + k sin a (x | y), restored per voxel
            # (reference two_stage.py:195-198, 239-244).
            k = 2 * np.pi * float(plan.fd) / float(plan.c)
            theta = plan.angles.astype(np.float64)
            kz, ku = k * (1 + np.cos(theta)), k * np.sin(theta)
            for name, axis, kk in (("x_phase", x, ku), ("z_phase", z, kz), ("y_phase", y, ku)):
                table = np.exp(1j * axis.astype(np.float64)[:, None] * kk[None, :]).astype(np.complex64)
                self._dev[name] = put(table, t.complex64)
        d, h = self._dev, self._host
        p = _lib.TwoStagePlan()
        p.ns, p.nc, p.na, p.nt = ns, nc, na, nt
        p.nx, p.nz, p.ny, p.npz = nx, nz, ny, len(plan.plane_z)
        p.n_planes, p.chunk, p.n_bands = plan.n_planes, self._chunk, len(bands)
        for field, key in (("band_depth_start", "depth_start"), ("band_depth_count", "depth_count"),
                           ("band_knot_first", "knot_first"), ("band_knot_count", "knot_count"),
                           ("band_radius", "radius"), ("knots_host", "knots")):
            setattr(p, field, h[key].ctypes.data)
        for field in ("x", "z", "y", "x_el", "y_el", "angles", "starts", "slopes", "knots", "x_centers", "x_plane",
                      "x_offset", "y_centers", "y_plane", "y_offset", "x_same", "y_same"):
            setattr(p, field, dev.ptr(d[field]))
        p.spline_upper, p.spline_inverse = dev.ptr(d["upper"]), dev.ptr(d["inverse"])
        for field in ("x_phase", "z_phase", "y_phase"):
            setattr(p, field, dev.ptr(d[field]) if baseband else None)
        p.c, p.fs, p.fd, p.dz = float(plan.c), float(plan.fs), float(plan.fd), float(plan.dz)
        p.baseband, p.stsw = int(bool(baseband)), int(self._stsw)
        self._struct = p
        lib = _lib.load()
        per = int(lib.rcabeam_two_stage_buffer_elements(ctypes.byref(p)))
        self._buffers = dev.empty((4 * per,), t.complex64)
        if self._stsw:
            nv = nx * nz * ny
            self._out = dev.empty((10, nv), f32)
            self._weights = dev.empty((nx, nz, ny), f32)
            self._power = dev.empty((nx, nz, ny), f32)
            self._peak = dev.empty((1,), t.int32)
        else:
            self._out = dev.empty((nx, nz, ny, nt), t.complex64)
        dev.synchronize()

    def _check_context(self) -> None:
        t = dev.torch()
        if t.cuda.current_device() != self._device.index or not dev.on_default_stream():
            raise ValueError("Workspace requires its construction device and the default CUDA stream")

    @property
    def plane_depths(self) -> np.ndarray:
        """A copy of the padded reference-depth coordinates."""
        return self._plan.plane_z.copy()

    @property
    def scratch_bytes(self) -> int:
        """Plane and spline scratch bytes, excluding raw input and the outputs."""
        return int(self._buffers.numel()) * 8

    def reconstruct(self, rc, cr):
        """Reconstruct contiguous complex64 device channels; the output is overwritten by the next call."""
        t = dev.torch()
        with self._lock:
            self._check_context()
            for raw in (rc, cr):
                if (not isinstance(raw, t.Tensor) or tuple(raw.shape) != self._shape or raw.dtype != t.complex64
                        or raw.device.type != "cuda"):
                    raise ValueError("Inputs must be matching complex64 CUDA ensembles")
                if not raw.is_contiguous() or raw.device.index != self._device.index:
                    raise ValueError("Inputs must be contiguous and on the workspace device")
            _lib.call("rcabeam_two_stage_reconstruct_device", ctypes.byref(self._struct), dev.ptr(rc), dev.ptr(cr),
                      dev.ptr(self._buffers), dev.ptr(self._out))
            if self._stsw:
                _lib.call("rcabeam_two_stage_stsw_finish_device", dev.ptr(self._out), dev.ptr(self._weights),
                          dev.ptr(self._power), self._weights.numel(), self._shape[3], dev.ptr(self._peak))
                result = self._weights, self._power
            else:
                result = self._out
            dev.synchronize()
            return result


class TwoStageStswWorkspace(TwoStageOpwWorkspace):
    """Two-subset baseband St-SW on the two-stage expansion (reference two_stage.py:399-420).

    `reconstruct` returns borrowed `(weights, power)`: weights normalised by
    their maximum and XDoppler power weighted by them, both `(nx, nz, ny)`.
    Ten real slow-time sums per voxel are accumulated over all frames.
    """

    _stsw = True

    def __init__(self, channel_shape, angles, t_start, grid, geom: RCAGeometry, *, n_planes=1,
                 samples_per_wavelength=4):
        if len(angles) < 2:
            raise ValueError("Two-subset St-SW requires at least two angles")
        super().__init__(channel_shape, angles, t_start, grid, geom, n_planes=n_planes,
                         samples_per_wavelength=samples_per_wavelength, baseband=True)
