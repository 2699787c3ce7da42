"""Low-rank + residual coding of 16-bit channel storage (mixed-precision OPW).

No reference counterpart: the reference's FP16 path rounds the channel IQ
directly (src/rcabeam/ensemble.cu:129-141), which leaves the rounding error of
the 0 dB tissue echoes on the -40 dB blood signal after clutter filtering.
Here the strongest temporal modes V (nt x k) of the channel Casorati matrix
(rows = samples x channels x angles of both apertures, columns = frames) are
split off first: each row c is stored as its residual e = c - (c V) V^H in
16 bits plus its k coefficients c V as 16-bit hi + lo pairs in 2k extra frame
columns.  The delay-and-sum of the coded rows (any 16-bit form of
`rcabeam_opw_grid_device`) is expanded back to the delay-and-sum of c by
`rcabeam_lowrank_expand_device`; only the residual carries 16-bit rounding.
"""

from __future__ import annotations

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd.ensemble import _RANGE_MESSAGE, _Geometry
from rcabeam_amd.sim import RCAGeometry

MAX_BASIS = 12  # rcabeam_pack_ensemble_lowrank_device: k <= 12
ROW_STEP = 4  # channel rows sampled for the basis: every 4th (sample, channel, angle) row
SUBSPACE_ITERATIONS = 12  # rounds of rcabeam_dominant_subspace_device for the basis


def channel_basis(rc, cr, k: int, row_step: int = ROW_STEP, iterations: int = SUBSPACE_ITERATIONS):
    """The k strongest temporal modes of the RC and CR channels together: complex128 (k, nt), row j = column j
    of V (orthonormal).

    The Gram is formed from every `row_step`-th row of both channel Casorati matrices (rows = (sample,
    channel, angle), complex64) with split bfloat16 (hi + lo) products on MFMA; its dominant k-dimensional
    subspace comes from `iterations` rounds of float64 subspace iteration on the GPU (one workgroup,
    rcabeam_dominant_subspace_device) rather than a full eigendecomposition.  Any orthonormal V codes the
    channels exactly up to the 16-bit rounding of the residual; V only has to catch the strong tissue
    modes, which every depth band carries and which stand well above the rest of the spectrum."""
    t = dev.torch()
    if not 1 <= k <= MAX_BASIS:
        raise ValueError(f"k must be in [1, {MAX_BASIS}]")
    if row_step < 1:
        raise ValueError("row_step must be positive")
    nt = rc.shape[-1]
    if k >= nt:
        raise ValueError("k must be smaller than the frame count")
    if k * nt > 4096:
        raise ValueError("k x frames must be at most 4096 (rcabeam_dominant_subspace_device)")
    if iterations < 1:
        raise ValueError("iterations must be positive")
    lib = _lib.load()
    stats = dev.empty((3,), t.int32)
    for i, a in enumerate((rc, cr)):
        _lib.call("rcabeam_iq_absmax_device", dev.ptr(a), 0, a.numel(), dev.ptr(stats), int(i > 0))
    total = rc.numel() // nt
    rows = (total + row_step - 1) // row_step
    nbytes = int(lib.rcabeam_gram_workspace_bytes(rows, nt))
    if nbytes == 0:
        raise ValueError("Frame count out of range for the GPU Gram")
    work = dev.empty((nbytes // 4,), t.float32)
    real = dev.empty((int(lib.rcabeam_gram_real_elements(nt)),), t.float64)
    for i, a in enumerate((rc, cr)):
        _lib.call("rcabeam_temporal_gram_sampled_device", dev.ptr(a), 0, rows, row_step, nt, dev.ptr(stats),
                  dev.ptr(real), dev.ptr(work), nbytes, int(i > 0))
    gram = dev.empty((nt, nt), t.complex128)
    _lib.call("rcabeam_gram_hermitian_device", dev.ptr(real), nt, 1.0 / (2 * rows), dev.ptr(gram))
    basis = dev.empty((k, nt), t.complex128)
    _lib.call("rcabeam_dominant_subspace_device", dev.ptr(gram), nt, k, iterations, dev.ptr(basis))
    return basis


def pack_channels_lowrank(raw, storage: str, basis, packed=None, range_error=None):
    """Code one complex64 device ensemble `(ns, nc, na, nt)` into the packed `(nc, na, ns, nt + 2k)` layout of
    16-bit words (int32 tensor).  Range errors raise `ValueError` as `pack_channels` does."""
    t = dev.torch()
    if storage not in ("float16", "bfloat16"):
        raise ValueError("Low-rank coding stores float16 or bfloat16 channels")
    ns, nc, na, nt = raw.shape
    k = basis.shape[0]
    if tuple(basis.shape) != (k, nt) or basis.dtype != t.complex128:
        raise ValueError("basis must be complex128 (k, frames)")
    if packed is None:
        packed = dev.empty((nc, na, ns, nt + 2 * k), t.int32)
    if range_error is None:
        range_error = dev.empty((1,), t.int32)
    range_error.zero_()
    _lib.call("rcabeam_pack_ensemble_lowrank_device", dev.ptr(raw), dev.ptr(packed), _lib.STORAGE[storage],
              dev.ptr(basis), k, dev.ptr(range_error), ns, nc, na, nt)
    if int(range_error.item()):
        raise ValueError(_RANGE_MESSAGE[storage])
    return packed


def beamform_lowrank(rc, cr, angles, t_start, grid, geom: RCAGeometry, *, storage: str = "float16", k: int = 4,
                     coefficients: int = _lib.COEF_16BIT, iq_storage: str = "float32"):
    """OPW IQ `(nv, nt)` (device tensor of `iq_storage`) and power `(nv, 2)` float32 of complex64 device
    channels `(ns, nc, na, nt)` through k-mode low-rank + 16-bit residual coding on a Cartesian grid."""
    t = dev.torch()
    dev.require_default_stream("The low-rank OPW requires the default stream")
    rc = dev.to_device(rc, t.complex64)
    cr = dev.to_device(cr, t.complex64)
    if rc.shape != cr.shape or rc.ndim != 4:
        raise ValueError("RC and CR channels must share the (samples, channels, angles, frames) shape")
    ns, nc, na, nt = rc.shape
    g = _Geometry(angles, t_start, grid, geom)
    basis = channel_basis(rc, cr, k)
    packed = [pack_channels_lowrank(x, storage, basis) for x in (rc, cr)]
    ntc = nt + 2 * k
    coded = dev.empty((g.nv, ntc), t.complex64)
    scratch = dev.empty((g.nv, 2), t.float32)
    nx, nz, ny = g.shape
    c, fs, fd, fn = g.constants
    _lib.call("rcabeam_opw_grid_device", dev.ptr(packed[0]), dev.ptr(packed[1]), _lib.STORAGE[storage], coefficients,
              dev.ptr(g.axes[0]), dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye),
              dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(scratch), dev.ptr(coded), 0, 0, ns, nc, na, ntc,
              c, fs, fd, fn)
    code = _lib.STORAGE[iq_storage]
    out = dev.empty((g.nv, nt), t.complex64 if code == 0 else t.int32)
    pd = dev.empty((g.nv, 2), t.float32)
    _lib.call("rcabeam_lowrank_expand_device", dev.ptr(coded), dev.ptr(basis), k, dev.ptr(out), code, dev.ptr(pd),
              g.nv, nt)
    return out, pd


__all__ = ["MAX_BASIS", "ROW_STEP", "beamform_lowrank", "channel_basis", "pack_channels_lowrank"]
