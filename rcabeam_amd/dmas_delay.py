"""Hybrid DMAS-CCF-ACF on the GPU: half-sample delay tables and axial splines.

Mirrors reference examples/_support/dmas_delay.py:61-154: `prepare_hybrid(shape, grid, geom,
angles, step, slab, cubic=False)` forms the geometry shared by successive ensembles on one
acquisition grid and returns `reconstruct(raw)`, which maps `[rc, cr]` channel IQ
`(samples, channels, angles, frames)` to the DMAS-CCF-ACF power `(nx, nz, ny)` float32.

* Angle planes: one baseband two-stage reference plane per aperture (four reference depths per
  wavelength) holds each angle's channel sum; a not-a-knot spline along depth gives the value at
  each voxel's mapped depth (`rcabeam_hybrid_planes_device`, `rcabeam_two_stage_spline_device`).
* Channel tables: per depth slab of `slab` depths, each channel's angle sum at the slab's
  steering anchor plus k * `step` samples (`rcabeam_hybrid_tables_device`); each voxel reads it
  at its excess receive delay, linearly or (`cubic`) with Catmull-Rom weights.
* `rcabeam_hybrid_finish_device` forms DMAS x CCF x ACF per voxel and frame and averages it.

The geometry follows the reference's float32 steps, including its fixed 2 us acquisition start
(dmas_delay.py:38-39, 100).  The tables and every sum run in HIP kernels; there is no CPU path.
"""

from __future__ import annotations

import numpy as np

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd.ensemble import pack_channels
from rcabeam_amd.sim import RCAGeometry
from rcabeam_amd.two_stage import _Plan, _spline_factors

_T0 = 2e-6  # acquisition start the prototype's geometry assumes


def _plane_geometry(shape, grid, geom: RCAGeometry, angles):
    """Per aperture `(prx (nv, nc, nk, 4), ptx (nk, na), tx (nz, nu, na, 4))` float32 and the knot
    count (dmas_delay.py:36-58)."""
    plan = _Plan(shape, angles, _T0, grid, geom, 4, 1, True)
    a, starts, knots = plan.angles, plan.starts, plan.plane_z
    x, z, y = plan.axes
    c, fs, fd = float(geom.c), float(geom.fs), float(geom.f_demod)
    wavenumber = 2 * np.pi * fd / c
    theta = a.astype(np.float64)
    kz, ku = wavenumber * (1 + np.cos(theta)), wavenumber * np.sin(theta)
    out = []
    for v, u, el, (centers, _, offset) in zip((x, y), (y, x), plan.elements, plan.layouts[::-1]):
        d = v[:, None, None] - el[:, None]
        r = np.sqrt(knots**2 + d**2)
        phase = 2 * np.pi * fd * (d**2 / ((r + knots) * c))
        aperture = np.ones_like(r) if geom.fnumber is None else (np.abs(d) <= knots / (2 * geom.fnumber))
        prx = np.stack((r * (fs / c), np.cos(phase), np.sin(phase), aperture.astype(np.float32)), axis=-1)
        ptx = ((knots[:, None] * np.cos(a) + centers[0] * np.sin(a)) / c - starts) * fs
        pos = (z[:, None, None] + offset[:, None] * plan.slopes - knots[0]) / np.float32(c / fd / 4)
        ph = z[:, None, None] * kz + u[:, None] * ku
        tx = np.stack((pos, np.cos(ph), np.sin(ph), np.zeros_like(ph)), axis=-1)
        out.append(tuple(np.ascontiguousarray(val, dtype=np.float32) for val in (prx, ptx, tx)))
    return out, len(knots)


def _slab_geometry(grid, geom: RCAGeometry, angles, step, slab):
    """Receive tables `(nz, nv, nc, 4)`, slab anchors `(n_slabs, nu, na, 4)` per aperture and the
    knots of each slab's channel tables (dmas_delay.py:66-105)."""
    x, z, y = (np.asarray(g) for g in grid)
    c, fs, fd = float(geom.c), float(geom.fs), float(geom.f_demod)
    nz = len(z)
    angles = np.asarray(angles)
    anchor_id = np.minimum((np.arange(nz) // slab) * slab + slab // 2, nz - 1)
    shift_z = z - z[anchor_id]
    common_shift = (z - z[(np.arange(nz) // slab) * slab]) * (2 * fs / c)
    receive_tables, anchors, sizes = [], [], []
    for v, u, el in ((x, y, geom.x_el), (y, x, geom.y_el)):
        dv = np.asarray(v, np.float32)[:, None] - np.asarray(el, np.float32)
        zz = np.asarray(z, np.float32)[:, None, None]
        receive = np.sqrt(zz**2 + dv**2) / c
        active = np.ones(receive.shape, bool) if geom.fnumber is None else np.abs(dv) <= zz / (2 * geom.fnumber)
        delta = np.where(active, (receive - zz / c) * fs + common_shift[:, None, None], -1).astype(np.float32)
        phase = (receive + shift_z[:, None, None] / c) * (2 * np.pi * fd)
        receive_tables.append(np.ascontiguousarray(np.stack(
            (delta, np.cos(phase).astype(np.float32), np.sin(phase).astype(np.float32), np.zeros_like(delta)),
            axis=-1)))
        sizes.append(delta.max(axis=(1, 2)).clip(0, None))
        za = z[anchor_id[::slab], None, None]
        steer = (za * np.cos(angles) + u[:, None] * np.sin(angles)) / c
        sample = (steer + za / c - _T0) * fs + (z[::slab, None, None] - za) * (2 * fs / c)
        ph = steer * (2 * np.pi * fd)
        anchors.append(np.ascontiguousarray(
            np.stack((sample, np.cos(ph), np.sin(ph), np.zeros_like(sample)), axis=-1), dtype=np.float32))
    counts = [int(np.ceil(max(s[first:first + slab].max() for s in sizes) / step)) + 2 for first in range(0, nz, slab)]
    return receive_tables, anchors, counts


def prepare_hybrid(shape, grid, geom: RCAGeometry, angles, step, slab, cubic=False):
    """Precompute the geometry shared by successive ensembles on this acquisition grid and return
    `reconstruct(raw)` (reference examples/_support/dmas_delay.py:61-154).  `step`: channel-table
    knot spacing in samples; `slab`: depths per channel table; `cubic`: Catmull-Rom table reads."""
    x, z, y = grid
    if np.any(np.diff(np.asarray(z)) < 0):
        raise ValueError("Hybrid depth slabs require a nondecreasing depth axis")
    if not (np.isfinite(step) and step > 0) or int(slab) < 1:
        raise ValueError("Hybrid tables need a positive step and slab")
    ns, nc, na, _ = (int(n) for n in shape)
    t = dev.torch()
    dev.device()
    planes_geometry, nk = _plane_geometry(shape, grid, geom, angles)
    receive_tables, anchors, counts = _slab_geometry(grid, geom, angles, float(step), int(slab))
    slab = int(slab)
    nx, nz, ny = len(x), len(z), len(y)
    upload = lambda a: dev.to_device(a, t.float32)  # noqa: E731
    geometry = [tuple(upload(a) for a in cfg) for cfg in planes_geometry]
    rx = [upload(a) for a in receive_tables]
    anchor = [upload(a) for a in anchors]
    upper, inverse = (upload(a) for a in _spline_factors(nk))
    lines = (nx, ny)  # receive lines: x for RC, y for CR
    transmit = (ny, nx)

    def reconstruct(raw):
        if len(raw) != 2:
            raise ValueError("raw must hold the RC and CR channel IQ")
        host = not all(dev.is_tensor(v) for v in raw)
        data = [dev.to_device(v, t.complex64) for v in raw]
        if any(tuple(v.shape[:3]) != (ns, nc, na) or v.ndim != 4 for v in data) or data[0].shape != data[1].shape:
            raise ValueError("raw must be two (samples, channels, angles, frames) arrays of the prepared shape")
        nt = int(data[0].shape[-1])
        packed = [pack_channels(v, "float32") for v in data]
        planes, seconds = [], []
        for cfg in range(2):
            plane = dev.empty((lines[cfg], nk, na, nt), t.complex64)
            second = dev.empty(tuple(plane.shape), t.complex64)
            prx, ptx, _ = geometry[cfg]
            _lib.call("rcabeam_hybrid_planes_device", dev.ptr(packed[cfg]), dev.ptr(prx), dev.ptr(ptx),
                      dev.ptr(plane), lines[cfg], nc, na, ns, nt, nk)
            _lib.call("rcabeam_two_stage_spline_device", dev.ptr(plane), dev.ptr(second), dev.ptr(upper),
                      dev.ptr(inverse), lines[cfg], nk, na, nt, nt)
            planes.append(plane)
            seconds.append(second)
        out = dev.empty((nx, nz, ny), t.float32)
        tables = [dev.empty((transmit[cfg] * nc * max(counts) * nt,), t.complex64) for cfg in range(2)]
        for first in range(0, nz, slab):
            nq = counts[first // slab]
            for cfg in range(2):
                _lib.call("rcabeam_hybrid_tables_device", dev.ptr(packed[cfg]), dev.ptr(anchor[cfg]),
                          dev.ptr(tables[cfg]), ns, nc, na, nt, transmit[cfg], first // slab, nq, float(step))
            _lib.call("rcabeam_hybrid_finish_device", dev.ptr(tables[0]), dev.ptr(tables[1]), dev.ptr(planes[0]),
                      dev.ptr(planes[1]), dev.ptr(seconds[0]), dev.ptr(seconds[1]), dev.ptr(rx[0]),
                      dev.ptr(geometry[0][2]), dev.ptr(rx[1]), dev.ptr(geometry[1][2]), dev.ptr(out), nx, nz, ny, nc,
                      na, nt, first, min(slab, nz - first), nq, nk, float(np.float32(1) / np.float32(step)),
                      int(bool(cubic)))
        return dev.host(out) if host else out

    return reconstruct
