"""Float64 restatement of the reference two-stage OPW approximation.

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`).

The reference's experimental two-stage path (`src/rcabeam/two_stage.py`,
`src/rcabeam/two_stage.cu`) beamforms each aperture onto a few dense
"reference planes" (lateral line x uniformly spaced reference depths, one
plane per elevation slab) and expands every voxel from them with a uniform
not-a-knot cubic spline along depth, after mapping the voxel's elevation
offset to a depth shift offset * tan(angle / 2).  Functions and the lines they
follow:

* `plane_layout`          two_stage.py:20-25 (equal-width slabs, nearest plane, offsets)
* `plan`                  two_stage.py:28-80 (validation, reference-depth axis padded for every mapping)
* `aperture_bands`        two_stage.py:83-102 (depth runs with one receive mask; knot slices with 3 guards)
* `unchanged_apertures`   two_stage.py:105-116 (lines whose channel set equals the previous band's)
* `second_derivatives`    two_stage.py:119-143; solved here as the full not-a-knot banded system
                          (rows 0 and n-1 are the not-a-knot conditions), not by the reference's
                          reduced Thomas recursion
* `two_stage_per_angle`   two_stage.py:146-192 and two_stage.cu:155-207 (per-angle, per-aperture
                          expanded signals; `baseband` removes and restores the closest-element
                          carrier, two_stage.py:168-172, 188-190, 195-198)
* `two_stage_stsw`        two_stage.cu:125-152, 198-239 and two_stage.py:384-392: two-subset St-SW
                          weights (normalised by their maximum) times XDoppler power of the expanded
                          per-angle signals, i.e. stsw.py:85-106 with k = 2
* `two_stage_opw`         sum of `two_stage_per_angle` over apertures and angles

Precision: the discrete plan (slab centres, reference-depth axis, bands, aperture masks) is
computed in float32 exactly as the reference defines it; delays, interpolation, the spline solve
and the voxel mapping are float64 from those float32 values.

Pinned in `tests/test_two_stage.py` by: scipy's not-a-knot `CubicSpline` (reference
tests/test_two_stage.py:34-42); exact agreement with the conventional delay-and-sum oracle at zero
steering (test_two_stage.py:45-56); the closest-element exact case and its slow-time phase
(test_two_stage.py:59-71); motion phase (test_two_stage.py:167-183); the aperture taken at the
output depth, with and without baseband (test_two_stage.py:192-224); channel-set flags against
brute-force sets (test_two_stage.py:102-109); and St-SW against the St-SW oracle on per-angle
signals obtained by zeroing all other angles (test_two_stage.py:227-259).
"""

from __future__ import annotations

import warnings

import numpy as np

from oracle.combiners import st_sw_weights
from oracle.das import delay_volumes, scan_coords


def plane_layout(axis, n_planes: int):
    """Slab centres, nearest centre per coordinate and offsets (two_stage.py:20-25)."""
    axis = np.asarray(axis, dtype=np.float32)
    lo, hi = float(axis.min()), float(axis.max())
    width = (hi - lo) / n_planes
    centers = np.array([lo + (k + 0.5) * width for k in range(n_planes)], dtype=np.float32)
    gap = np.abs(axis[:, None] - centers[None, :])
    nearest = np.argmin(gap, axis=1).astype(np.int32)
    return centers, nearest, (axis - centers[nearest]).astype(np.float32)


def plan(shape, angles, t_start, grid, x_el, y_el, c, fs, fd, fnumber, *, samples_per_wavelength, n_planes,
         baseband=False):
    """Reference-plane geometry (two_stage.py:28-80); raises ValueError like the reference."""
    if isinstance(n_planes, (bool, np.bool_)) or int(n_planes) < 1:
        raise ValueError("Require at least one reference plane")
    if len(shape) != 4 or any(int(n) < 1 for n in shape) or shape[0] < 2:
        raise ValueError("Expected nonempty (samples, channels, angles, frames), samples >=2")
    axes = tuple(np.array(v, dtype=np.float32) for v in grid)
    angles = np.array(angles, dtype=np.float32)
    elements = (np.array(x_el, dtype=np.float32), np.array(y_el, dtype=np.float32))
    if len(angles) != shape[2] or any(len(e) != shape[1] for e in elements):
        raise ValueError("Geometry must match channel and angle dimensions")
    starts = np.array(np.broadcast_to(np.asarray(t_start, dtype=np.float32), angles.shape))
    fn = np.float32(-1.0 if fnumber is None else fnumber)
    spw = float(samples_per_wavelength)
    if not (np.isfinite(spw) and np.isfinite([c, fs, fd, fn]).all()):
        raise ValueError("Geometry and sampling must be finite")
    if min(c, fs, fd) <= 0 or spw < 2 or (fnumber is not None and fn <= 0):
        raise ValueError("Require positive c/fs/carrier/F-number and at least two axial samples per wavelength")
    if not baseband and spw <= 4 * (1 + fs / (2 * fd)):
        warnings.warn("Axial undersampling experiment", UserWarning, stacklevel=2)
    x, z, y = axes
    if np.any(z <= 0) or np.any(np.abs(angles) >= np.pi / 2):
        raise ValueError("Require positive depths and steering angles strictly within +/-pi/2")
    layouts = (plane_layout(x, n_planes), plane_layout(y, n_planes))
    slopes = np.tan(angles / np.float32(2)).astype(np.float32)
    shifts = np.concatenate([(off[:, None] * slopes[None, :]).ravel() for _, _, off in layouts]).astype(np.float32)
    dz = np.float32(c / fd / spw)
    z0 = np.float32(z.min() + shifts.min() - np.float32(2) * dz)
    end = np.float32(z.max() + shifts.max() + np.float32(2) * dz)
    if z0 <= 0:
        raise ValueError("Mapped reference planes reach nonpositive depths")
    npz = int(np.ceil(np.float32(end - z0) / dz)) + 1
    plane_z = (z0 + np.arange(npz, dtype=np.float32) * dz).astype(np.float32)
    return dict(shape=tuple(int(n) for n in shape), angles=angles, starts=starts, axes=axes, elements=elements,
                layouts=layouts, slopes=slopes, shifts=shifts, dz=dz, plane_z=plane_z,
                constants=(np.float32(c), np.float32(fs), np.float32(fd), fn))


def aperture_bands(p):
    """`(depth slice, knot slice)` per run of depths sharing every receive mask (two_stage.py:83-102)."""
    x, z, y = p["axes"]
    fn = p["constants"][3]
    pz, shifts = p["plane_z"], p["shifts"]
    if fn <= 0:
        return [(slice(0, len(z)), slice(0, len(pz)))]
    pairs = np.sort(np.concatenate([np.abs(a[:, None] - e[None, :]).ravel()
                                    for a, e in zip((x, y), p["elements"])]).astype(np.float32))
    radius = (z / (np.float32(2) * fn)).astype(np.float32)
    counts = np.searchsorted(pairs, radius, side="right")
    cuts = [0] + [i + 1 for i in range(len(z) - 1) if counts[i + 1] != counts[i]] + [len(z)]
    out = []
    for first, stop in zip(cuts[:-1], cuts[1:]):
        depths = z[first:stop]
        lo = max(0, int(np.searchsorted(pz, np.float32(depths.min() + shifts.min()))) - 3)
        hi = min(len(pz), int(np.searchsorted(pz, np.float32(depths.max() + shifts.max()), side="right")) + 3)
        out.append((slice(first, stop), slice(lo, hi)))
    return out


def unchanged_apertures(lateral, elements, band_depths, fnumber):
    """uint8 `(bands, lines)`: 1 where a line's active channel set equals the previous band's
    (two_stage.py:105-116), from explicit channel sets."""
    lateral = np.asarray(lateral, np.float32)
    elements = np.asarray(elements, np.float32)
    radius = (np.asarray(band_depths, np.float32) / (np.float32(2) * np.float32(fnumber))).astype(np.float32)
    active = np.abs(lateral[None, :, None] - elements[None, None, :]) <= radius[:, None, None]
    out = np.zeros((len(radius), len(lateral)), np.uint8)
    out[1:] = (active[1:] == active[:-1]).all(axis=-1)
    return out


def second_derivatives(planes, axis: int = 1):
    """Not-a-knot cubic-spline second derivatives times the knot step squared, along `axis`
    (two_stage.py:119-143), from the full banded system in float64."""
    p = np.moveaxis(np.asarray(planes, dtype=np.complex128), axis, 0)
    n = p.shape[0]
    if n < 4:
        raise ValueError("a not-a-knot spline needs at least four knots")
    a = np.zeros((n, n))
    rhs = np.zeros_like(p)
    a[0, :3] = (1, -2, 1)
    a[-1, -3:] = (1, -2, 1)
    for i in range(1, n - 1):
        a[i, i - 1:i + 2] = (1, 4, 1)
        rhs[i] = 6 * (p[i - 1] - 2 * p[i] + p[i + 1])
    m = np.linalg.solve(a, rhs.reshape(n, -1)).reshape(p.shape)
    return np.moveaxis(m, 0, axis)


def _wavenumbers(angles, c, fd):
    """Closest-element carrier wavenumbers k (1 + cos a) and k sin a (two_stage.py:195-198)."""
    k = 2 * np.pi * float(fd) / float(c)
    th = np.asarray(angles, np.float64)
    return k * (1 + np.cos(th)), k * np.sin(th)


def two_stage_per_angle(rc, cr, angles, t_start, grid, x_el, y_el, c, fs, fd, fnumber, *,
                        samples_per_wavelength=8, n_planes=1, baseband=False, float32_delays=False):
    """Per-angle expanded signals `(rc_part, cr_part)`, each `(nx, nz, ny, na, nt)` complex128.

    `float32_delays` reproduces the reference's float32 geometry: plane delays in float32 (its plane
    grids are float32 arrays, so `_delay_numpy` runs in float32 and evaluates exp(2j pi fd tau) in
    complex64, two_stage.py:163-165) and float32 knot positions (two_stage.py:178).  The default is
    the float64 answer, which the GPU kernels follow."""
    rc = np.asarray(rc)
    cr = np.asarray(cr)
    if rc.shape != cr.shape:
        raise ValueError("RC and CR channel shapes must match")
    p = plan(rc.shape, angles, t_start, grid, x_el, y_el, c, fs, fd, fnumber,
             samples_per_wavelength=samples_per_wavelength, n_planes=n_planes, baseband=baseband)
    x, z, y = p["axes"]
    c32, fs32, fd32, fn = p["constants"]
    fnum = None if fn <= 0 else float(fn)
    na, nt = rc.shape[2], rc.shape[3]
    dz = float(p["dz"])
    kz, ku = _wavenumbers(p["angles"], c32, fd32)
    parts = [np.zeros((len(x), len(z), len(y), na, nt), np.complex128) for _ in range(2)]
    for band, knot_slice in aperture_bands(p):
        knots = p["plane_z"][knot_slice]
        zb = z[band].astype(np.float64)
        for config, raw, layout, elev, elements in ((0, rc, p["layouts"][1], y, p["elements"][0]),
                                                    (1, cr, p["layouts"][0], x, p["elements"][1])):
            centers, nearest, offset = layout
            lateral = x if config == 0 else y
            plane_grid = (lateral, knots, centers) if config == 0 else (centers, knots, lateral)
            per_angle, _, _ = delay_volumes(raw, scan_coords(plane_grid), elements, p["angles"], p["starts"],
                                            c32, fs32, fd32, fnum, "RC" if config == 0 else "CR",
                                            aperture_depth=z[band.start], float32_delays=float32_delays)
            if config == 0:
                planes = per_angle.reshape(len(lateral), len(knots), len(centers), na, nt)
            else:
                planes = per_angle.reshape(len(centers), len(knots), len(lateral), na, nt).transpose(2, 1, 0, 3, 4)
            if baseband:
                phase = (knots.astype(np.float64)[:, None, None] * kz[None, None, :]
                         + centers.astype(np.float64)[None, :, None] * ku[None, None, :])
                planes = planes * np.exp(-1j * phase)[None, ..., None]
            second = second_derivatives(planes, axis=1)
            for m in range(na):
                if float32_delays:
                    # float32 knot position, as the reference forms it (two_stage.py:178)
                    pos = ((z[band, None] + offset[None, :] * p["slopes"][m] - knots[0]) / p["dz"]).astype(np.float64)
                else:
                    pos = (zb[:, None] + offset.astype(np.float64)[None, :] * float(p["slopes"][m])
                           - float(knots[0])) / dz
                i = np.floor(pos).astype(np.int64)
                if i.min() < 0 or i.max() >= len(knots) - 1:
                    raise AssertionError("voxel mapped outside its knot slice")
                w = pos - i
                a = 1 - w
                # planes[:, i, nearest, m]: (lateral, band depth, elevation, frames)
                lo = planes[:, i, nearest[None, :], m]
                hi = planes[:, i + 1, nearest[None, :], m]
                slo = second[:, i, nearest[None, :], m]
                shi = second[:, i + 1, nearest[None, :], m]
                wa, ww = a[None, ..., None], w[None, ..., None]
                value = wa * lo + ww * hi + ((wa**3 - wa) * slo + (ww**3 - ww) * shi) / 6
                if baseband:
                    ph = zb[:, None] * kz[m] + elev.astype(np.float64)[None, :] * ku[m]
                    value = value * np.exp(1j * ph)[None, ..., None]
                if config == 0:
                    parts[0][:, band, :, m] = value
                else:
                    parts[1][:, band, :, m] = value.transpose(2, 1, 0, 3)
    return parts[0], parts[1]


def two_stage_opw(rc, cr, angles, t_start, grid, x_el, y_el, c, fs, fd, fnumber, **options):
    """Two-stage OPW IQ `(nx, nz, ny, nt)` complex128 (two_stage.py:146-192)."""
    r, q = two_stage_per_angle(rc, cr, angles, t_start, grid, x_el, y_el, c, fs, fd, fnumber, **options)
    return (r + q).sum(axis=-2)


def two_stage_stsw(rc, cr, angles, t_start, grid, x_el, y_el, c, fs, fd, fnumber, *, n_planes=1,
                   samples_per_wavelength=4):
    """`(weights, power)` of the two-subset baseband St-SW (two_stage.py:384-420, two_stage.cu:125-239)."""
    if len(angles) < 2:
        raise ValueError("Two-subset St-SW requires at least two angles")
    r, q = two_stage_per_angle(rc, cr, angles, t_start, grid, x_el, y_el, c, fs, fd, fnumber,
                               samples_per_wavelength=samples_per_wavelength, n_planes=n_planes, baseband=True)
    na = r.shape[-2]
    weights = st_sw_weights(np.concatenate([r, q], axis=-2), np.arange(na), np.arange(na, 2 * na), k=2)
    xd = np.abs((r.sum(axis=-2) * np.conj(q.sum(axis=-2))).mean(axis=-1))
    return weights, weights * xd
