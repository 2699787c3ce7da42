"""Single-frame power volumes straight from channel data, and combiners of per-angle volumes, on the GPU.

Mirrors reference src/rcabeam/fused.py:62-298 (`opw_pd_from_channels`,
`xdoppler_pd_from_channels`, `rc_fmas_pd_from_channels`,
`fast_pd_from_channels`, `dmas_ccf_acf_from_channels`), the GPU branches of
src/rcabeam/opw.py:24-51 (`opw`), src/rcabeam/xdoppler.py:65-117
(`xdoppler_pd`) and src/rcabeam/fmas.py:54-89 (`rc_fmas_pd`).  Every function
runs the HIP kernels behind the per-frame C ABI (include/rcabeam.h); `use_cuda`
is accepted for signature compatibility and must stay True (this package has no
CPU path).  Angle index sets that are not one contiguous run are gathered on
the device before the span kernels run.
"""

from __future__ import annotations

from typing import Literal

import numpy as np

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd.sim import RCAGeometry, _fnumber, _scan_coords

_NO_CPU = "rcabeam_amd has no CPU path; use_cuda must be True"


def _channel_args(iq_rc, iq_cr, angles, t_start, grid, geom: RCAGeometry):
    """Device arguments of the fused per-frame kernels (reference fused.py:15-38)."""
    t = dev.torch()
    rc = dev.to_device(iq_rc, t.complex64)
    cr = dev.to_device(iq_cr, t.complex64)
    if rc.ndim != 3 or tuple(rc.shape) != tuple(cr.shape):
        raise ValueError("RC and CR channel IQ must be matching (samples, channels, angles) arrays")
    ns, nc, na = rc.shape
    if ns < 2 or len(geom.x_el) != nc or len(geom.y_el) != nc or len(angles) != na:
        raise ValueError("Expected at least two samples and matching channel/angle dimensions")
    starts = np.ascontiguousarray(np.broadcast_to(np.asarray(t_start, dtype=np.float32), (na,)))
    scan = dev.to_device(_scan_coords(grid), t.float32)
    vectors = [dev.to_device(np.asarray(v, dtype=np.float32), t.float32) for v in (geom.x_el, geom.y_el, angles)]
    st = dev.to_device(starts, t.float32)
    ptrs = [dev.ptr(rc), dev.ptr(cr), dev.ptr(scan), *(dev.ptr(v) for v in vectors), dev.ptr(st)]
    sizes = (ns, nc, na, scan.shape[0])
    constants = (float(geom.c), float(geom.fs), float(geom.f_demod), _fnumber(geom))
    keep = (rc, cr, scan, vectors, st)
    return ptrs, sizes, constants, keep


def _single(name, iq_rc, iq_cr, angles, t_start, grid, geom, use_cuda):
    if not use_cuda:
        raise ValueError(_NO_CPU)
    t = dev.torch()
    ptrs, sizes, constants, keep = _channel_args(iq_rc, iq_cr, angles, t_start, grid, geom)
    out = dev.empty((sizes[3],), t.float32)
    _lib.call(name, *ptrs, dev.ptr(out), *sizes, *constants)
    del keep
    return dev.host(out).reshape(tuple(len(g) for g in grid))


def opw_pd_from_channels(iq_rc, iq_cr, angles, t_start, grid, geom: RCAGeometry, *, use_cuda: bool = True):
    """OPW power `(nx, nz, ny)` of one frame (reference fused.py:169-209)."""
    return _single("rcabeam_opw_pd_from_channels_device", iq_rc, iq_cr, angles, t_start, grid, geom, use_cuda)


def xdoppler_pd_from_channels(iq_rc, iq_cr, angles, t_start, grid, geom: RCAGeometry, *, use_cuda: bool = True):
    """XDoppler power `(nx, nz, ny)` of one frame (reference fused.py:255-298)."""
    return _single("rcabeam_xdoppler_pd_from_channels_device", iq_rc, iq_cr, angles, t_start, grid, geom, use_cuda)


def rc_fmas_pd_from_channels(iq_rc, iq_cr, angles, t_start, grid, geom: RCAGeometry, *, use_cuda: bool = True):
    """RC-FMAS power `(nx, nz, ny)` of one frame (reference fused.py:212-252)."""
    return _single("rcabeam_rc_fmas_pd_from_channels_device", iq_rc, iq_cr, angles, t_start, grid, geom, use_cuda)


def dmas_ccf_acf_from_channels(iq_rc, iq_cr, angles, t_start, grid, geom: RCAGeometry, *, use_cuda: bool = True):
    """DMAS x CCF x ACF `(nx, nz, ny)` of one frame (reference fused.py:124-166)."""
    return _single("rcabeam_dmas_ccf_acf_from_channels_device", iq_rc, iq_cr, angles, t_start, grid, geom, use_cuda)


def fast_pd_from_channels(iq_rc, iq_cr, angles, t_start, grid, geom: RCAGeometry, *, use_cuda: bool = True):
    """OPW, XDoppler and RC-FMAS power volumes of one frame in one pass (reference fused.py:62-113)."""
    if not use_cuda:
        raise ValueError(_NO_CPU)
    t = dev.torch()
    ptrs, sizes, constants, keep = _channel_args(iq_rc, iq_cr, angles, t_start, grid, geom)
    outs = [dev.empty((sizes[3],), t.float32) for _ in range(3)]
    _lib.call("rcabeam_fast_pd_from_channels_device", *ptrs, *(dev.ptr(o) for o in outs), *sizes, *constants)
    del keep
    shape = tuple(len(g) for g in grid)
    return tuple(dev.host(o).reshape(shape) for o in outs)


def opw(iq, *, use_cuda: bool = True):
    """Coherent angle sum of per-angle volumes `(..., n_angles, n_frames)` -> `(..., n_frames)` complex64
    (reference opw.py:24-51)."""
    if not use_cuda:
        raise ValueError(_NO_CPU)
    t = dev.torch()
    host = not dev.is_tensor(iq)
    arr = dev.to_device(iq, t.complex64)
    if arr.ndim < 2:
        raise ValueError("iq must have at least angle and frame axes")
    na, nt = arr.shape[-2], arr.shape[-1]
    nv = arr.numel() // (na * nt)
    out = dev.empty((*arr.shape[:-2], nt), t.complex64)
    if nv:
        _lib.call("rcabeam_opw_device", dev.ptr(arr), dev.ptr(out), nv, na, nt)
    return dev.host(out) if host else out


def _span(indices) -> tuple[int, int] | None:
    """(start, count) of a contiguous run of angle indices, else None (reference xdoppler.py:54-62)."""
    idx = np.asarray(indices)
    if idx.size == 0:
        return None
    start = int(idx[0])
    return (start, int(idx.size)) if np.array_equal(idx, np.arange(start, start + idx.size)) else None


def _per_angle(iq, rc_idx, cr_idx):
    """Device volumes and the (rc, cr) angle spans; non-contiguous index sets are gathered on the device."""
    t = dev.torch()
    arr = dev.to_device(iq, t.complex64)
    if arr.ndim < 2:
        raise ValueError("iq must have at least angle and frame axes")
    numbers = np.arange(arr.shape[-2])
    rc, cr = numbers[rc_idx], numbers[cr_idx]
    rs, cs = _span(rc), _span(cr)
    if rs is None or cs is None:
        order = dev.to_device(np.concatenate([np.atleast_1d(rc), np.atleast_1d(cr)]).astype(np.int64), t.int64)
        arr = arr.index_select(arr.ndim - 2, order).contiguous()
        rs, cs = (0, int(np.size(rc))), (int(np.size(rc)), int(np.size(cr)))
    return arr, rs, cs


def xdoppler_pd(iq, rc_idx, cr_idx, *, output: Literal["abs", "real", "complex"] = "abs", use_cuda: bool = True):
    """XDoppler power `(...)` of per-angle volumes (reference xdoppler.py:65-117): `abs`, `real` or `complex`
    of mean_t (sum over rc_idx) conj(sum over cr_idx)."""
    if not use_cuda:
        raise ValueError(_NO_CPU)
    if output not in ("abs", "real", "complex"):
        raise ValueError("output must be 'abs', 'real', or 'complex'")
    t = dev.torch()
    host = not dev.is_tensor(iq)
    arr, rs, cs = _per_angle(iq, rc_idx, cr_idx)
    na, nt = arr.shape[-2], arr.shape[-1]
    lead = tuple(arr.shape[:-2])
    nv = arr.numel() // (na * nt)
    if output == "abs":
        out = dev.empty(lead, t.float32)
        if nv:
            _lib.call("rcabeam_xdoppler_pd_device", dev.ptr(arr), dev.ptr(out), nv, na, nt, *rs, *cs)
    else:
        mean = dev.empty((*lead, 2), t.float32)
        if nv:
            _lib.call("rcabeam_xdoppler_mean_device", dev.ptr(arr), dev.ptr(mean), nv, na, nt, *rs, *cs)
        out = mean[..., 0].contiguous() if output == "real" else t.view_as_complex(mean)
    return dev.host(out) if host else out


def rc_fmas_pd(iq, rc_idx, cr_idx, *, use_cuda: bool = True):
    """Mean RC-FMAS power `(...)` of per-angle volumes (reference fmas.py:54-89)."""
    if not use_cuda:
        raise ValueError(_NO_CPU)
    t = dev.torch()
    host = not dev.is_tensor(iq)
    arr, rs, cs = _per_angle(iq, rc_idx, cr_idx)
    na, nt = arr.shape[-2], arr.shape[-1]
    nv = arr.numel() // (na * nt)
    out = dev.empty(tuple(arr.shape[:-2]), t.float32)
    if nv:
        _lib.call("rcabeam_rc_fmas_pd_device", dev.ptr(arr), dev.ptr(out), nv, na, nt, *rs, *cs)
    return dev.host(out) if host else out
