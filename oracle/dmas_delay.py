"""Hybrid DMAS-CCF-ACF approximation (delay tables + axial splines), CPU restatement.

TEST INFRASTRUCTURE ONLY: imported by tests, `__graft_entry__.smoke()` and
`bench.py`'s CPU baseline leg, never by the product path.

Follows reference examples/_support/dmas_delay.py:20-154 and
examples/_support/dmas_delay.cu:21-151.  The DMAS term needs each channel's
angle-summed signal and the ACF needs each angle's channel-summed signal; the
hybrid forms both from tables instead of full delay-and-sum:

* channel tables C (dmas_delay.cu:21-38): per depth slab, transmit line u,
  channel and knot k, the angle sum of samples taken at the slab's steering
  anchor plus k * step samples, rotated by the anchor's carrier phase;
* angle planes A (dmas_delay.cu:96-114): two-stage reference planes (one
  baseband plane per aperture, four reference depths per wavelength) with a
  not-a-knot spline along depth, looked up at the voxel's mapped depth;
* finish (dmas_delay.cu:116-151): per voxel and frame, the receive excess
  delay of each active channel indexes C (linear, or Catmull-Rom inside the
  table), the rotated values feed the signed-root DMAS sums and the CCF
  magnitudes, the A lookups feed the per-angle ACF sums, and the power
  DMAS x CCF x ACF is averaged over frames.

Geometry tables are formed with the reference's float32 steps (the GPU
kernels consume float32 tables); interpolation, sums and the power are float64.
The plane geometry assumes an acquisition start of 2 us, as the reference's
prototype does (dmas_delay.py:38-39, 100).

Pinning (tests/test_dmas_delay.py): on the reference's broadside case
(tests/test_dmas_delay_benchmark.py:11-52: four elements, three zero angles,
depths on carrier-wavelength knots, affine IQ in sample index) the hybrid
equals exact DMAS-CCF-ACF of `oracle.das` + `oracle.combiners` within the
reference's 3e-4, for the linear and the cubic table; DMAS-CCF-ACF is
positively homogeneous of degree 1 in a common channel gain; a plausible
mistake (conjugated anchor phase, swapped apertures, a dropped excess term)
breaks the broadside equality.
"""

from __future__ import annotations

import numpy as np

from oracle import two_stage

T0 = 2e-6  # acquisition start the prototype's geometry assumes (dmas_delay.py:38-39, 100)


def _f32(v):
    return np.asarray(v, dtype=np.float32)


def plane_geometry(shape, grid, x_el, y_el, angles, c, fs, fd, fnumber):
    """Receive / transmit terms of the angle planes and their voxel lookups (dmas_delay.py:36-58).

    Returns per aperture (RC receives along x, CR along y) `(prx, ptx, tx)`:
    prx `(nv, nc, nk, 4)` = (receive samples r fs / c, cos, sin of the baseband excess phase,
    aperture flag); ptx `(nk, na)` transmit samples of the slab centre; tx `(nz, nu, na, 4)` =
    (knot position, cos, sin of the residual carrier phase, 0); and the knot count."""
    p = two_stage.plan(shape, angles, T0, grid, x_el, y_el, c, fs, fd, fnumber, samples_per_wavelength=4,
                       n_planes=1, baseband=True)
    a, starts = p["angles"], p["starts"]
    x, z, y = p["axes"]
    knots = p["plane_z"]
    kz, ku = two_stage._wavenumbers(a, c, fd)
    out = []
    for v, u, el, (centers, _, offset) in zip((x, y), (y, x), p["elements"], p["layouts"][::-1]):
        d = v[:, None, None] - el[:, None]
        r = np.sqrt(knots**2 + d**2)
        excess = d**2 / ((r + knots) * c)
        phase = 2 * np.pi * fd * excess
        aperture = np.ones_like(r) if fnumber is None else (np.abs(d) <= knots / (2 * fnumber)).astype(np.float32)
        prx = np.stack((r * (fs / c), np.cos(phase), np.sin(phase), aperture), axis=-1).astype(np.float32)
        ptx = (((knots[:, None] * np.cos(a) + centers[0] * np.sin(a)) / c - starts) * fs).astype(np.float32)
        dz = np.float32(c / fd / 4)
        pos = (z[:, None, None] + offset[:, None] * p["slopes"] - knots[0]) / dz
        ph = z[:, None, None] * kz + u[:, None] * ku
        tx = np.stack((pos, np.cos(ph), np.sin(ph), np.zeros_like(ph)), axis=-1).astype(np.float32)
        out.append((prx, ptx, tx))
    return out, len(knots)


def hybrid_geometry(shape, grid, x_el, y_el, angles, c, fs, fd, fnumber, step, slab):
    """Receive tables, slab anchors and table lengths (dmas_delay.py:61-105).

    Returns `(rx, anchors, counts)`: rx per aperture `(nz, nv, nc, 4)` = (excess receive delay in
    samples plus the slab's common two-way shift, or -1 outside the aperture; cos, sin of the
    receive carrier phase referred to the slab anchor; 0); anchors per aperture
    `(n_slabs, nu, na, 4)` = (anchor sample, cos, sin of the steering phase, 0); counts: knots per
    slab's channel tables."""
    x, z, y = (np.asarray(g) for g in grid)
    if np.any(np.diff(z) < 0):
        raise ValueError("Hybrid depth slabs require a nondecreasing depth axis")
    nz = len(z)
    angles = np.asarray(angles)
    anchor_id = np.minimum((np.arange(nz) // slab) * slab + slab // 2, nz - 1)
    dz = z - z[anchor_id]
    common_shift = (z - z[(np.arange(nz) // slab) * slab]) * (2 * fs / c)
    rx, anchors, sizes = [], [], []
    for v, u, el in ((x, y, x_el), (y, x, y_el)):
        dv = _f32(v)[:, None] - _f32(el)
        zz = _f32(z)[:, None, None]
        receive = np.sqrt(zz**2 + dv**2) / c
        active = np.ones(receive.shape, bool) if fnumber is None else np.abs(dv) <= zz / (2 * fnumber)
        delta = (receive - zz / c) * fs + common_shift[:, None, None]
        phase = (receive + dz[:, None, None] / c) * (2 * np.pi * fd)
        delta = np.where(active, delta, -1).astype(np.float32)
        rx.append(np.stack((delta, np.cos(phase).astype(np.float32), np.sin(phase).astype(np.float32),
                            np.zeros_like(delta)), axis=-1))
        sizes.append(delta.max(axis=(1, 2)).clip(0, None))
        za = z[anchor_id[::slab], None, None]
        steer = (za * np.cos(angles) + u[:, None] * np.sin(angles)) / c
        sample = (steer + za / c - T0) * fs + (z[::slab, None, None] - za) * (2 * fs / c)
        ph = steer * (2 * np.pi * fd)
        anchors.append(np.stack((sample, np.cos(ph), np.sin(ph), np.zeros_like(sample)), axis=-1).astype(np.float32))
    counts = [int(np.ceil(max(s[first:first + slab].max() for s in sizes) / step)) + 2 for first in range(0, nz, slab)]
    return rx, anchors, counts


def _sample(trace, pos):
    """Linear interpolation of `trace (ns, nt)` at sample positions `pos` (any shape); zero outside
    the closed window [0, ns - 1] (dmas_delay.cu:11-16)."""
    ns = trace.shape[0]
    pos = np.asarray(pos, np.float64)
    ok = (pos >= 0) & (pos <= ns - 1)
    p = np.where(ok, pos, 0.0)
    i = np.floor(p).astype(np.int64)
    w = (p - i)[..., None]
    v = trace[i] * (1 - w) + trace[np.minimum(i + 1, ns - 1)] * w
    return np.where(ok[..., None], v, 0)


def _rotate(v, g):
    """v e^{i phi} with (cos, sin) = g[..., 1:3] (dmas_delay.cu:4-6)."""
    return v * (g[..., 1].astype(np.float64) + 1j * g[..., 2].astype(np.float64))[..., None]


def angle_planes(packed, prx, ptx):
    """A[(v, k, a), t] = sum over aperture channels of the rotated sample at r + ptx[k, a]
    (dmas_delay.cu:96-114); packed `(nc, na, ns, nt)`."""
    nc, na, ns, nt = packed.shape
    nv, _, nk, _ = prx.shape
    out = np.zeros((nv, nk, na, nt), np.complex128)
    for ch in range(nc):
        g = prx[:, ch]  # (nv, nk, 4)
        for a in range(na):
            val = _rotate(_sample(packed[ch, a].astype(np.complex128), g[..., 0].astype(np.float64) + ptx[None, :, a]), g)
            out[:, :, a] += np.where(g[..., 3:4] != 0, val, 0)
    return out


def channel_tables(packed, anchors, z_slab, nq, step):
    """C[(u, ch, k), t] = sum over angles of the anchor-rotated sample at anchor + k step
    (dmas_delay.cu:21-38) for one slab."""
    nc, na, ns, nt = packed.shape
    nu = anchors.shape[1]
    out = np.zeros((nu, nc, nq, nt), np.complex128)
    k = np.arange(nq, dtype=np.float32) * np.float32(step)
    for ch in range(nc):
        for a in range(na):
            b = anchors[z_slab, :, a]  # (nu, 4)
            pos = (b[:, 0:1] + k[None, :]).astype(np.float32).astype(np.float64)
            val = _sample(packed[ch, a].astype(np.complex128), pos)  # (nu, nq, nt)
            out[:, ch] += _rotate(val, b[:, None, :])
    return out


def _table(rows, position, cubic):
    """Value of table rows `(n, nq, nt)` at per-row positions `(n,)` (dmas_delay.cu:40-61)."""
    n, nq, _ = rows.shape
    idx = np.arange(n)
    i = np.minimum(np.floor(position).astype(np.int64), nq - 2)
    w = (position - i)[:, None]
    lin = rows[idx, i] * (1 - w) + rows[idx, i + 1] * w
    if not cubic:
        return lin
    ic = np.floor(position).astype(np.int64)
    inner = (ic != 0) & (ic + 2 < nq)
    j = np.clip(ic, 1, nq - 3)
    wc = (position - j)[:, None]
    p_1, p0, p1, p2 = (rows[idx, j + s] for s in (-1, 0, 1, 2))
    value = p0 * (1 - wc) + p1 * wc
    left, right = p_1 + p1 - 2 * p0, p0 + p2 - 2 * p1
    cub = value - 0.5 * wc * (1 - wc) * (left * (1 - wc) + right * wc)
    return np.where(inner[:, None], cub, lin)


def _lookup(plane, second, position):
    """Not-a-knot spline value of planes `(nk, nt)` rows at `position` (two_stage.cu:113-123)."""
    nk = plane.shape[0]
    i = int(np.floor(position))
    if i < 0 or i >= nk - 1:
        return np.zeros(plane.shape[1], np.complex128)
    w = float(position) - i
    a = 1 - w
    return (plane[i] * a + plane[i + 1] * w + second[i] * ((a**3 - a) / 6) + second[i + 1] * ((w**3 - w) / 6))


def hybrid_dmas(rc, cr, angles, grid, x_el, y_el, c, fs, fd, fnumber, *, step, slab, cubic=False):
    """Hybrid DMAS-CCF-ACF power `(nx, nz, ny)` of channels `(ns, nc, na, nt)` per aperture
    (dmas_delay.py:61-154)."""
    rc, cr = (np.asarray(v) for v in (rc, cr))
    if rc.ndim == 3:
        rc, cr = rc[..., None], cr[..., None]
    ns, nc, na, nt = rc.shape
    x, z, y = (np.asarray(g) for g in grid)
    nx, nz, ny = len(x), len(z), len(y)
    plan, nk = plane_geometry(rc.shape, grid, x_el, y_el, angles, c, fs, fd, fnumber)
    rx, anchors, counts = hybrid_geometry(rc.shape, grid, x_el, y_el, angles, c, fs, fd, fnumber, step, slab)
    packed = [np.ascontiguousarray(v.transpose(1, 2, 0, 3)) for v in (rc, cr)]
    planes, seconds = [], []
    for cfg in range(2):
        a = angle_planes(packed[cfg], plan[cfg][0], plan[cfg][1])
        planes.append(a)
        seconds.append(two_stage.second_derivatives(a, axis=1))
    out = np.zeros((nx, nz, ny))
    istep = np.float32(1) / np.float32(step)
    for first in range(0, nz, slab):
        nq = counts[first // slab]
        tables = [channel_tables(packed[cfg], anchors[cfg], first // slab, nq, step) for cfg in range(2)]
        for iz in range(first, min(first + slab, nz)):
            for ix in range(nx):
                for iy in range(ny):
                    roots = np.zeros(nt, np.complex128)
                    magnitude = np.zeros(nt)
                    active = 0
                    coherent = [np.zeros(nt, np.complex128) for _ in range(2)]
                    energy = [np.zeros(nt) for _ in range(2)]
                    for cfg in range(2):
                        v, u = (iy, ix) if cfg else (ix, iy)
                        r = rx[cfg][iz, v]  # (nc, 4)
                        on = np.flatnonzero(r[:, 0] >= 0)
                        if len(on):
                            pos = (r[on, 0] * istep).astype(np.float32).astype(np.float64)
                            vals = _rotate(_table(tables[cfg][u, on], pos, cubic), r[on])
                            m = np.abs(vals)
                            roots += np.sum(np.where(m > 0, vals / np.sqrt(np.where(m > 0, m, 1)), 0), axis=0)
                            magnitude += m.sum(axis=0)
                            active += len(on)
                        b = plan[cfg][2][iz, u]  # (na, 4)
                        for a in range(na):
                            val = _lookup(planes[cfg][v, :, a], seconds[cfg][v, :, a], float(b[a, 0])) \
                                * (float(b[a, 1]) + 1j * float(b[a, 2]))
                            coherent[cfg] += val
                            energy[cfg] += np.abs(val) ** 2
                    dm = np.abs(roots) ** 2 - magnitude
                    rr = np.abs(coherent[0]) ** 2 / na**2
                    qq = np.abs(coherent[1]) ** 2 / na**2
                    var = np.maximum(0, energy[0] / na - rr) + np.maximum(0, energy[1] / na - qq)
                    num = np.sqrt(rr) * np.sqrt(qq)
                    den = rr + qq + 0.5 * var - num
                    ok = (dm > 0) & (magnitude > 0) & (active > 0) & (den > 0)
                    power = np.where(ok, dm * (dm / max(active, 1) / np.where(magnitude > 0, magnitude, 1))
                                     * num / np.where(den > 0, den, 1), 0)
                    out[ix, iz, iy] = power.sum() / nt
    return out
