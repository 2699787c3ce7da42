"""Frame-batched RCA reconstruction on MI355X.

Public functions mirror reference src/rcabeam/ensemble.py: `opw_voxel_order`
(:24-49), `ensemble_pd_from_channels` (:133-218), `beamform_ensemble`
(:221-270), `beamform_apertures` (:273-287), with the same validation messages
(:66-84), broadcasting of per-angle start times (:82) and St-SW normalisation
(:127-129).  The staging that the reference binding does in C++
(src/rcabeam/bindings/python.cpp:495-571: upload, pack, voxel tiles,
download) happens here over the C ABI with torch-owned device buffers.

Inputs may be NumPy arrays (uploaded, results returned as NumPy, as in the
reference) or torch tensors already on the GPU (results stay on the GPU).
"""

from __future__ import annotations

from time import perf_counter
from typing import Literal, TypedDict

import numpy as np
from numpy.typing import NDArray

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd.sim import RCAGeometry, _scan_coords

Method = Literal["opw", "xdoppler", "rc_fmas", "dmas", "st_sw", "st_sw_baseband"]
IQStorage = Literal["float32", "float16", "bfloat16"]
_METHODS = {"opw": 0, "xdoppler": 1, "rc_fmas": 2, "dmas": 3, "st_sw": 4, "apertures": 5}
_RANGE_MESSAGE = {
    "float16": "FP16 IQ requires finite components within +/-65504; rescale inputs explicitly",
    "bfloat16": "BF16 IQ requires finite components within the bfloat16 range; rescale inputs explicitly",
}
# Shape limits of rcabeam_opw_grid_device (include/rcabeam_mi355x.h).
GRID_MAX_ANGLES = 16
GRID_MAX_FRAMES = 208
# Bound on the St-SW correlation workspace per launch (voxel tiles beyond it).
_STSW_WORKSPACE_BYTES = 1 << 28
# Device bytes the angle-by-angle combiners may hold at once (per-angle IQ + accumulators).
_COMBINE_BYTES = 96 << 30
# Channels per DMAS accumulator pass.
_DMAS_BATCH = 4


class EnsembleTiming(TypedDict):
    """Host wall time for completed raw RC/CR uploads, excluding allocation."""

    raw_upload_seconds: float


def opw_voxel_order(shape: tuple[int, int, int]) -> NDArray[np.intp]:
    """Voxel ids in 2 (x) by 4 (y) lateral patches, tails kept (ensemble.py:24-49).

    Scatter values computed in this order into these ids to restore `(x, z, y)`.
    """
    if min(shape) < 1:
        raise ValueError("Grid dimensions must be positive")
    nx, nz, ny = shape
    ids = np.arange(nx * nz * ny, dtype=np.intp).reshape(shape)
    ex, ey = nx - nx % 2, ny - ny % 4
    patches = ids[:ex, :, :ey].reshape(ex // 2, 2, nz, ey // 4, 4).transpose(0, 2, 3, 1, 4).ravel()
    return np.concatenate((patches, ids[:ex, :, ey:].ravel(), ids[ex:, :, :].ravel()))


def _validate(rc, cr, angles, grid, geom: RCAGeometry, method: str, k: int, return_signal: bool, iq_storage: str):
    if method not in _METHODS or (method == "apertures" and not return_signal):
        raise ValueError(f"Unknown method: {method}")
    if iq_storage not in ("float32", "float16", "bfloat16") or (iq_storage != "float32" and method != "opw"):
        raise ValueError("IQ storage must be float32, or float16/bfloat16 for OPW only")
    shape = tuple(rc.shape)
    if len(shape) != 4 or tuple(cr.shape) != shape or any(n == 0 for n in shape) or shape[0] < 2:
        raise ValueError("RC/CR must have matching nonempty (samples, channels, angles, frames) shapes")
    if len(angles) != shape[2] or len(geom.x_el) != shape[1] or len(geom.y_el) != shape[1]:
        raise ValueError("Geometry must match the channel/angle dimensions")
    if any(np.ndim(axis) != 1 or not len(axis) or not np.isfinite(axis).all() for axis in grid):
        raise ValueError("Grid axes must be nonempty finite coordinate vectors")
    if not np.isfinite([geom.c, geom.fs, geom.f_demod]).all() or geom.c <= 0 or geom.fs <= 0:
        raise ValueError("Sound speed and sampling frequency must be finite and positive")
    if geom.fnumber is not None and (not np.isfinite(geom.fnumber) or geom.fnumber <= 0):
        raise ValueError("F-number must be positive or None")
    if method == "st_sw" and (k < 2 or k > 4 or len(angles) < k):
        raise ValueError("St-SW requires 2 <= k <= 4 and at least k angles")


def pack_channels(raw, storage: str, packed=None, range_error=None):
    """Pack one complex64 device ensemble `(ns, nc, na, nt)` into the frame-inner layout.

    Returns the packed device buffer (complex64, or int32 words holding two
    16-bit components).  FP16/BF16 range errors raise `ValueError`.
    """
    t = dev.torch()
    ns, nc, na, nt = raw.shape
    if packed is None:
        packed = dev.empty((nc, na, ns, nt), t.complex64 if storage == "float32" else t.int32)
    if storage == "float32":
        _lib.call("rcabeam_pack_ensemble_device", dev.ptr(raw), dev.ptr(packed), ns, nc, na, nt)
        return packed
    if range_error is None:
        range_error = dev.empty((1,), t.int32)
    range_error.zero_()
    name = "rcabeam_pack_ensemble_fp16_device" if storage == "float16" else "rcabeam_pack_ensemble_bf16_device"
    _lib.call(name, dev.ptr(raw), dev.ptr(packed), dev.ptr(range_error), ns, nc, na, nt)
    if int(range_error.item()):
        raise ValueError(_RANGE_MESSAGE[storage])
    return packed


class _Geometry:
    """Float32 device copies of everything the ensemble entry points read."""

    def __init__(self, angles, t_start, grid, geom: RCAGeometry, order=None):
        t = dev.torch()
        na = len(angles)
        starts = np.ascontiguousarray(np.broadcast_to(np.asarray(t_start, dtype=np.float32), (na,)))
        vectors = [np.asarray(v, dtype=np.float32) for v in (geom.x_el, geom.y_el, angles)]
        if not all(np.isfinite(v).all() for v in (*vectors, starts)):
            raise ValueError("Geometry and sample starts must be finite")
        scan = _scan_coords(grid)
        self.canonical = order is None  # scan follows the grid's own (x, z, y) order
        if order is not None:
            scan = np.ascontiguousarray(scan[order])
        self.nv = len(scan)
        self.shape = tuple(len(axis) for axis in grid)
        self.axes = tuple(dev.to_device(np.asarray(axis, dtype=np.float32), t.float32) for axis in grid)
        self.scan = dev.to_device(scan, t.float32)
        self.xe, self.ye, self.angles = (dev.to_device(v, t.float32) for v in vectors)
        self.starts = dev.to_device(starts, t.float32)
        self.constants = (float(geom.c), float(geom.fs), float(geom.f_demod), -1.0 if geom.fnumber is None else float(geom.fnumber))
        # The reference ABI requires a non-null geometry workspace; it is unused here.
        self.geometry_workspace = dev.empty((16,), t.uint8)


def run_ensemble(packed_rc, packed_cr, g: _Geometry, shape4, method: int, k: int, storage: str, pd, signal,
                 weights=None, first: int = 0, count: int | None = None, signal_storage: int = 0):
    """Enqueue one reconstruction of voxels [first, first + count) over the C ABI.

    OPW and separate-aperture IQ of a whole grid run rcabeam_opw_grid_device
    (8 x 8 lateral tiles; the same taps as the voxel-list path); XDoppler of a
    whole grid runs its separate-aperture form followed by
    rcabeam_xdoppler_from_apertures_device.
    Otherwise float32 channels run rcabeam_ensemble_device and 16-bit channels
    (OPW only) rcabeam_opw_mixed_device, with float32-accurate coefficients.
    """
    t = dev.torch()
    ns, nc, na, nt = shape4
    count = g.nv - first if count is None else count
    c, fs, fd, fn = g.constants
    whole_grid = first == 0 and count == g.nv and g.canonical
    grid_shape_ok = whole_grid and na <= GRID_MAX_ANGLES and nt <= GRID_MAX_FRAMES

    def grid_call(grid_method, grid_pd, grid_signal, grid_signal_storage):
        nx, nz, ny = g.shape
        _lib.call(
            "rcabeam_opw_grid_device", dev.ptr(packed_rc), dev.ptr(packed_cr), _lib.STORAGE[storage],
            _lib.COEF_FLOAT32, dev.ptr(g.axes[0]), dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny,
            dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(grid_pd),
            dev.ptr(grid_signal), grid_signal_storage, grid_method, ns, nc, na, nt, c, fs, fd, fn,
        )

    if grid_shape_ok and method in (_METHODS["opw"], _METHODS["apertures"]):
        grid_call(method, pd, signal, signal_storage)
        return
    if grid_shape_ok and method == _METHODS["xdoppler"] and storage == "float32":
        # XDoppler of a whole grid: separate-aperture IQ on the grid kernel, then the cross product.
        planes = dev.empty((2 * g.nv, nt), t.complex64)
        grid_call(_METHODS["apertures"], pd, planes, 0)
        _lib.call("rcabeam_xdoppler_from_apertures_device", dev.ptr(planes), g.nv, nt, dev.ptr(pd), dev.ptr(signal))
        return
    scan = g.scan[first : first + count]
    pd_view = pd[first : first + count]
    planes = 2 if method == _METHODS["apertures"] else 1
    sig_view = None if signal is None else signal[first * planes : (first + count) * planes]
    common = (dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(pd_view), dev.ptr(sig_view))
    if storage == "float32":
        workspace = None
        w_view = None
        if method == _METHODS["st_sw"]:
            workspace = dev.empty((count * k * k * nt,), t.complex64)
            w_view = weights[first : first + count]
        _lib.call(
            "rcabeam_ensemble_device",
            dev.ptr(packed_rc), dev.ptr(packed_cr), dev.ptr(scan), *common, dev.ptr(w_view), dev.ptr(workspace),
            dev.ptr(g.geometry_workspace), ns, nc, na, nt, count, c, fs, fd, fn, method, k,
        )
    elif storage == "float16" and signal_storage == 0:
        # The reference's FP16 entry point (ensemble.cu:798-809): float32 interpolation and accumulation.
        _lib.call(
            "rcabeam_opw_fp16_device",
            dev.ptr(packed_rc), dev.ptr(packed_cr), dev.ptr(scan), *common, dev.ptr(g.geometry_workspace),
            ns, nc, na, nt, count, c, fs, fd, fn,
        )
    else:
        # 16-bit channels (OPW only): float32-accurate coefficients, float32 accumulation.
        _lib.call(
            "rcabeam_opw_mixed_device",
            dev.ptr(packed_rc), dev.ptr(packed_cr), dev.ptr(scan), *common, _lib.STORAGE[storage], signal_storage,
            _lib.COEF_FLOAT32, ns, nc, na, nt, count, c, fs, fd, fn,
        )


def _per_angle_ok(method: str, storage: str, na: int, nt: int, k: int) -> bool:
    """RC-FMAS and St-SW of a whole grid run angle by angle on the grid kernel."""
    if storage != "float32" or nt > GRID_MAX_FRAMES:
        return False
    return method == "rc_fmas" or (method == "st_sw" and 2 <= k <= 4 and na >= k)


def _slab_lines(shape, bytes_per_voxel: int) -> int:
    """Lines of x per slab so that the combiners' per-voxel buffers fit the device-memory budget
    (a multiple of the 8-wide lateral tile when the grid needs more than one slab)."""
    nx, nz, ny = shape
    budget = min(_COMBINE_BYTES, dev.torch().cuda.mem_get_info()[0] // 2)
    lines = min(nx, max(1, budget // (bytes_per_voxel * nz * ny)))
    if lines < nx and lines >= 8:
        lines -= lines % 8
    return lines


def run_per_angle(raw_rc, raw_cr, g: _Geometry, shape4, method: int, k: int, pd, signal, weights):
    """RC-FMAS / St-SW of a whole grid, angle by angle (reference ensemble.cu:236-315, 459-590).

    The channels are packed angle-major, (angles, channels, samples, frames), so each angle is a
    valid one-angle packed ensemble.  For each x-slab of the grid (sized to a device-memory
    budget, a multiple of the 8-wide lateral tile): one separate-aperture grid launch per angle,
    each folded into the combiner's accumulator, then the reduction over frames.
    """
    t = dev.torch()
    ns, nc, na, nt = shape4
    nx, nz, ny = g.shape
    c, fs, fd, fn = g.constants
    lib = _lib.load()
    fields = int(lib.rcabeam_angle_accumulator_fields(method, k))
    by_angle = [x.permute(2, 1, 0, 3).contiguous() for x in (raw_rc, raw_cr)]
    # Angles folded per accumulator pass: k for St-SW (each subset once per pass), 4 for RC-FMAS.
    batch = k if method == _METHODS["st_sw"] else 4
    nxs = _slab_lines(g.shape, nt * (batch * 2 * 8 + 4 * fields))
    slab = nxs * nz * ny
    planes = dev.empty((batch, 2 * slab, nt), t.complex64)
    acc = dev.empty((slab * fields * nt,), t.float32)
    scratch = dev.empty((slab, 2), t.float32)
    for x0 in range(0, nx, nxs):
        x1 = min(nx, x0 + nxs)
        nvs = (x1 - x0) * nz * ny
        v0 = x0 * nz * ny
        for a0 in range(0, na, batch):
            count = min(batch, na - a0)
            for b in range(count):
                a = a0 + b
                _lib.call(
                    "rcabeam_opw_grid_device", dev.ptr(by_angle[0][a]), dev.ptr(by_angle[1][a]),
                    _lib.STORAGE["float32"], _lib.COEF_FLOAT32, dev.ptr(g.axes[0][x0:x1]), dev.ptr(g.axes[1]),
                    dev.ptr(g.axes[2]), x1 - x0, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles[a : a + 1]),
                    dev.ptr(g.starts[a : a + 1]), dev.ptr(scratch), dev.ptr(planes[b]), 0, _METHODS["apertures"], ns,
                    nc, 1, nt, c, fs, fd, fn,
                )
            _lib.call("rcabeam_angle_accumulate_device", dev.ptr(planes), count, 2 * slab * nt, nvs, nt, method, k,
                      a0, dev.ptr(acc))
        _lib.call(
            "rcabeam_angle_combine_finish_device", dev.ptr(acc), nvs, nt, na, method, k, dev.ptr(pd[v0 : v0 + nvs]),
            dev.ptr(None if signal is None else signal[v0 : v0 + nvs]),
            dev.ptr(None if weights is None else weights[v0 : v0 + nvs]),
        )


def run_dmas_grid(raw_rc, raw_cr, g: _Geometry, shape4, pd):
    """DMAS-CCF-ACF of a whole grid from grid-kernel launches (reference dmas.py:32-85).

    In the packed (channels, angles, samples, frames) layout channel i of each aperture is itself a
    valid one-channel ensemble: one separate-aperture grid launch per channel gives both
    apertures' angle-summed channel signals (the DMAS and CCF terms), and one launch per angle on
    angle-major channels gives the per-angle signals (the ACF).  x-slabs bound the accumulator.
    """
    t = dev.torch()
    ns, nc, na, nt = shape4
    nx, nz, ny = g.shape
    c, fs, fd, fn = g.constants
    lib = _lib.load()
    fields = int(lib.rcabeam_dmas_accumulator_fields())
    packed = [pack_channels(x, "float32") for x in (raw_rc, raw_cr)]
    by_angle = [x.permute(2, 1, 0, 3).contiguous() for x in (raw_rc, raw_cr)]
    batch = _DMAS_BATCH  # channels whose outputs are folded in one pass over the accumulator
    nxs = _slab_lines(g.shape, nt * (batch * 2 * 8 + 4 * fields))
    slab = nxs * nz * ny
    planes = dev.empty((batch, 2 * slab, nt), t.complex64)
    acc = dev.empty((slab * fields * nt,), t.float32)
    scratch = dev.empty((slab, 2), t.float32)
    plane_elems = na * ns * nt

    def grid(rc_ptr, cr_ptr, x0, x1, xe, ye, angles, starts, channels, angle_count, out):
        _lib.call(
            "rcabeam_opw_grid_device", rc_ptr, cr_ptr, _lib.STORAGE["float32"], _lib.COEF_FLOAT32,
            dev.ptr(g.axes[0][x0:x1]), dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), x1 - x0, nz, ny, xe, ye, angles,
            starts, dev.ptr(scratch), dev.ptr(out), 0, _METHODS["apertures"], ns, channels, angle_count, nt, c, fs,
            fd, fn,
        )

    for x0 in range(0, nx, nxs):
        x1 = min(nx, x0 + nxs)
        nvs = (x1 - x0) * nz * ny
        v0 = x0 * nz * ny
        for i0 in range(0, nc, batch):
            count = min(batch, nc - i0)
            for b in range(count):
                offset = (i0 + b) * plane_elems * 8
                grid(dev.ptr(packed[0]) + offset, dev.ptr(packed[1]) + offset, x0, x1,
                     dev.ptr(g.xe[i0 + b : i0 + b + 1]), dev.ptr(g.ye[i0 + b : i0 + b + 1]), dev.ptr(g.angles),
                     dev.ptr(g.starts), 1, na, planes[b])
            _lib.call("rcabeam_dmas_accumulate_device", dev.ptr(planes), count, 2 * slab * nt, nvs, nt, 1,
                      int(i0 == 0), dev.ptr(acc))
        for a in range(na):
            grid(dev.ptr(by_angle[0][a]), dev.ptr(by_angle[1][a]), x0, x1, dev.ptr(g.xe), dev.ptr(g.ye),
                 dev.ptr(g.angles[a : a + 1]), dev.ptr(g.starts[a : a + 1]), nc, 1, planes[0])
            _lib.call("rcabeam_dmas_accumulate_device", dev.ptr(planes), 1, 2 * slab * nt, nvs, nt, 0, int(a == 0),
                      dev.ptr(acc))
        _lib.call("rcabeam_dmas_finish_device", dev.ptr(acc), dev.ptr(g.axes[0][x0:x1]), dev.ptr(g.axes[1]),
                  dev.ptr(g.axes[2]), x1 - x0, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), nc, na, nt, fn,
                  dev.ptr(pd[v0 : v0 + nvs]))


def _run(rc, cr, angles, t_start, grid, geom, method, k, return_signal, timings, iq_storage):
    """Validate, upload, pack, reconstruct and download (ensemble.py:52-130)."""
    _validate(rc, cr, angles, grid, geom, method, k, return_signal, iq_storage)
    t = dev.torch()
    host_input = not (dev.is_tensor(rc) and dev.is_tensor(cr))
    dev.device()
    _lib.load()
    shape = (len(grid[0]), len(grid[1]), len(grid[2]))
    g = _Geometry(angles, t_start, grid, geom)
    ns, nc, na, nt = tuple(rc.shape)
    nv = g.nv
    code = _METHODS[method]
    pd = dev.empty((nv, 2), t.float32)
    weights = dev.empty((nv,), t.float32) if method == "st_sw" else None
    planes = 2 if method == "apertures" else 1
    signal = dev.empty((nv * planes, nt), t.complex64) if return_signal else None
    if (host_input and method == "opw" and iq_storage == "float32" and na <= GRID_MAX_ANGLES
            and nt <= GRID_MAX_FRAMES and nt % 2 == 0):
        opw_signal = signal if signal is not None else dev.empty((nv, nt), t.complex64)
        upload = _opw_host_overlapped(rc, cr, g, (ns, nc, na, nt), pd, opw_signal)
    else:
        dev.synchronize()
        begin = perf_counter()
        raw_rc = dev.to_device(rc, t.complex64)
        raw_cr = dev.to_device(cr, t.complex64)
        dev.synchronize()
        upload = perf_counter() - begin if host_input else 0.0
        if _per_angle_ok(method, iq_storage, na, nt, k):
            run_per_angle(raw_rc, raw_cr, g, (ns, nc, na, nt), code, k, pd, signal, weights)
            del raw_rc, raw_cr
        elif method == "dmas" and iq_storage == "float32" and na <= GRID_MAX_ANGLES and nt <= GRID_MAX_FRAMES:
            run_dmas_grid(raw_rc, raw_cr, g, (ns, nc, na, nt), pd)
            del raw_rc, raw_cr
        else:
            _run_packed(raw_rc, raw_cr, g, (ns, nc, na, nt), method, code, k, iq_storage, pd, signal, weights)
    dev.synchronize()
    if timings is not None:
        timings["raw_upload_seconds"] = upload
    to_host = dev.host if host_input else (lambda x: x)
    if return_signal:
        signal_shape = (*shape, 2, nt) if method == "apertures" else (*shape, nt)
        return to_host(pd[:, 0].reshape(shape)), to_host(signal.reshape(signal_shape))
    power = pd[:, 0].clone()
    if method == "st_sw":
        peak = weights.max()
        power = power * (weights / peak) if float(peak) > 0 else power * 0
    return to_host(power.reshape(shape)), None


def _opw_host_overlapped(rc, cr, g: _Geometry, shape4, pd, signal) -> float:
    """OPW of host channels with the CR upload overlapping the RC delay-and-sum (no reference counterpart:
    the reference uploads both apertures before reconstructing, bindings/python.cpp:527-531).

    RC: staged upload, pack, grid pass over the RC channels into `signal`; CR: staged upload on the
    library's copy stream while that pass runs, pack, grid pass over the CR channels added to `signal`
    with the power of the total into `pd`.  Returns the host wall time of the two uploads."""
    t = dev.torch()
    ns, nc, na, nt = shape4
    hosts = [np.ascontiguousarray(x, dtype=np.complex64) for x in (rc, cr)]
    raw = [dev.empty(shape4, t.complex64) for _ in range(2)]
    packed = [dev.empty((nc, na, ns, nt), t.complex64) for _ in range(2)]
    threads = dev.upload_threads()
    nx, nz, ny = g.shape
    c, fs, fd, fn = g.constants
    dev.synchronize()  # from here on no queued work touches raw[1]
    seconds = 0.0
    for aperture in range(2):
        begin = perf_counter()
        name = "rcabeam_upload_host" if aperture == 0 else "rcabeam_upload_host_concurrent"
        _lib.call(name, hosts[aperture].ctypes.data, dev.ptr(raw[aperture]), hosts[aperture].nbytes, threads)
        seconds += perf_counter() - begin
        pack_channels(raw[aperture], "float32", packed=packed[aperture])
        _lib.call(
            "rcabeam_opw_grid_aperture_device", dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.axes[0]),
            dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
            dev.ptr(g.starts), dev.ptr(pd), dev.ptr(signal), aperture, aperture, ns, nc, na, nt, c, fs, fd, fn,
        )
    return seconds


def _run_packed(raw_rc, raw_cr, g, shape4, method, code, k, iq_storage, pd, signal, weights):
    """Pack and reconstruct through rcabeam_ensemble_device / the grid kernel (every other case)."""
    ns, nc, na, nt = shape4
    nv = g.nv
    packed_rc = pack_channels(raw_rc, iq_storage)
    packed_cr = pack_channels(raw_cr, iq_storage)
    del raw_rc, raw_cr
    if method == "st_sw":
        tile = max(1, _STSW_WORKSPACE_BYTES // (k * k * nt * 8))
        for first in range(0, nv, tile):
            run_ensemble(packed_rc, packed_cr, g, (ns, nc, na, nt), code, k, iq_storage, pd, signal, weights,
                         first, min(tile, nv - first))
    else:
        run_ensemble(packed_rc, packed_cr, g, (ns, nc, na, nt), code, k, iq_storage, pd, signal, weights)


def ensemble_pd_from_channels(rc, cr, angles, t_start, grid, geom: RCAGeometry, *, method: Method = "opw",
                              k: int = 2, timings: EnsembleTiming | None = None,
                              iq_storage: IQStorage = "float32"):
    """Reconstruct a channel ensemble and reduce it to a `(nx, nz, ny)` float32 volume.

    OPW / RC-FMAS: mean squared magnitude; XDoppler: |mean cross-product|;
    St-SW: XDoppler weighted by normalised ensemble correlations; DMAS: mean
    real DMAS-CCF-ACF response.  Unfiltered (reference ensemble.py:133-218).
    `st_sw_baseband` approximates the per-angle signals with the two-stage
    expansion (one baseband reference plane per aperture, four reference depths
    per wavelength; `rcabeam_amd.two_stage`) and needs k=2, float32 storage.
    `iq_storage` "float16"/"bfloat16" packs channels at 16 bits (OPW only);
    arithmetic stays float32.
    """
    if method == "st_sw_baseband":
        return _stsw_baseband(rc, cr, angles, t_start, grid, geom, k, timings, iq_storage)
    return _run(rc, cr, angles, t_start, grid, geom, method, k, False, timings, iq_storage)[0]


def _stsw_baseband(rc, cr, angles, t_start, grid, geom, k, timings, iq_storage):
    """Two-subset St-SW on the two-stage baseband approximation (reference ensemble.py:198-217):
    one reference plane per aperture at four reference depths per wavelength."""
    from rcabeam_amd.two_stage import TwoStageStswWorkspace

    if k != 2 or iq_storage != "float32":
        raise ValueError("Baseband St-SW requires k=2 and float32 IQ storage")
    if tuple(rc.shape) != tuple(cr.shape):
        raise ValueError("RC and CR channel shapes must match")
    t = dev.torch()
    workspace = TwoStageStswWorkspace(tuple(rc.shape), angles, t_start, grid, geom)
    host_input = not (dev.is_tensor(rc) and dev.is_tensor(cr))
    begin = perf_counter()
    raw = [dev.to_device(np.ascontiguousarray(x) if not dev.is_tensor(x) else x, t.complex64) for x in (rc, cr)]
    if timings is not None:
        dev.synchronize()
        timings["raw_upload_seconds"] = perf_counter() - begin if host_input else 0.0
    _, power = workspace.reconstruct(*raw)
    return dev.host(power) if host_input else power.clone()


def beamform_ensemble(rc, cr, angles, t_start, grid, geom: RCAGeometry, *,
                      method: Literal["opw", "xdoppler", "rc_fmas"] = "opw", timings: EnsembleTiming | None = None,
                      iq_storage: IQStorage = "float32"):
    """Complex64 `(nx, nz, ny, frames)` OPW IQ or XDoppler/RC-FMAS slow-time signal (ensemble.py:221-270)."""
    if method not in ("opw", "xdoppler", "rc_fmas"):
        raise ValueError("Only OPW, XDoppler and RC-FMAS have complex signal outputs")
    return _run(rc, cr, angles, t_start, grid, geom, method, 2, True, timings, iq_storage)[1]


def filtered_xdoppler_pd(rc, cr, angles, t_start, grid, geom: RCAGeometry, *, rank_start, rank_stop=None,
                         precision: str = "complex128"):
    """Clutter-filtered XDoppler power `(nx, nz, ny)` float32, the recipe of reference README.md:33:
    separately compounded RC and CR IQ (`beamform_apertures`), one temporal basis shared by both apertures
    (the Gram of both Casorati matrices, as the reference's RC/CR tuple input to `temporal_basis`,
    gpu_clutter.py:122-127), the singular-vector interval [rank_start, rank_stop) kept in both, then
    |mean_t RC conj(CR)|.  Everything stays on the GPU: the separate-aperture grid kernel, one Gram over the
    (voxel, aperture) rows, the float64 EVD, one filter pass and `rcabeam_xdoppler_from_apertures_device`."""
    from rcabeam_amd import clutter

    _validate(rc, cr, angles, grid, geom, "apertures", 2, True, "float32")
    t = dev.torch()
    host_input = not (dev.is_tensor(rc) and dev.is_tensor(cr))
    shape = (len(grid[0]), len(grid[1]), len(grid[2]))
    g = _Geometry(angles, t_start, grid, geom)
    ns, nc, na, nt = tuple(rc.shape)
    packed = [pack_channels(dev.to_device(x, t.complex64), "float32") for x in (rc, cr)]
    planes = dev.empty((2 * g.nv, nt), t.complex64)  # (voxel, aperture) rows
    scratch = dev.empty((g.nv, 2), t.float32)
    run_ensemble(packed[0], packed[1], g, (ns, nc, na, nt), _METHODS["apertures"], 2, "float32", scratch, planes)
    del packed
    _, basis = clutter.temporal_basis(planes, precision=precision)
    filtered = clutter.apply_temporal_filter(planes, basis, rank_start=rank_start, rank_stop=rank_stop)
    del planes
    pd = dev.empty((g.nv, 2), t.float32)
    _lib.call("rcabeam_xdoppler_from_apertures_device", dev.ptr(filtered), g.nv, nt, dev.ptr(pd), None)
    power = pd[:, 0].reshape(shape)
    return dev.host(power) if host_input else power


def beamform_apertures(rc, cr, angles, t_start, grid, geom: RCAGeometry):
    """Separately compounded RC and CR IQ, each `(nx, nz, ny, frames)` (ensemble.py:273-287)."""
    _, iq = _run(rc, cr, angles, t_start, grid, geom, "apertures", 2, True, None, "float32")
    return iq[..., 0, :], iq[..., 1, :]
