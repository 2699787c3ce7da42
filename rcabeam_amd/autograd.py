"""Differentiable OPW beamforming: forward and hand-written adjoint HIP kernels.

`opw_iq(rc, cr, angles, t_start, grid, geom, sound_speed=...)` returns the
OPW IQ volume `(nx, nz, ny, frames)` (complex64, on the GPU) as an autograd
node.  Its backward pass is `rcabeam_opw_adjoint_grid_deterministic_device`:
the transpose of the delay-and-sum applied to the IQ gradient (fixed-point
accumulation, bit-identical from run to run), giving the gradients of a
real loss with respect to the complex64 RC/CR channel IQ and the scalar sound
speed.  Tap weights and phases are recomputed in the backward pass from the
geometry (nothing but the packed channels is kept between the passes), frame
chunk by frame chunk.  The reference has no differentiable path; the forward
values are those of `beamform_ensemble` (reference src/rcabeam/ensemble.py:221).

Gradient convention (PyTorch's, for a real loss L of complex z = x + iy):
dL/dx + i dL/dy.
"""

from __future__ import annotations

import numpy as np

from rcabeam_amd import _device as dev
from rcabeam_amd import _lib
from rcabeam_amd.ensemble import GRID_MAX_ANGLES, _Geometry, pack_channels
from rcabeam_amd.sim import RCAGeometry


def _function():
    t = dev.torch()

    class OpwFunction(t.autograd.Function):
        """IQ = OPW(rc, cr; c) with the adjoint kernel as backward."""

        @staticmethod
        def forward(ctx, rc, cr, sound_speed, g: _Geometry):
            ns, nc, na, nt = rc.shape
            c = float(sound_speed.detach().item())
            _, fs, fd, fn = g.constants
            packed = [pack_channels(x.detach().contiguous(), "float32") for x in (rc, cr)]
            nx, nz, ny = g.shape
            pd = dev.empty((g.nv, 2), t.float32)
            iq = dev.empty((g.nv, nt), t.complex64)
            _lib.call(
                "rcabeam_opw_grid_device", dev.ptr(packed[0]), dev.ptr(packed[1]), 0, _lib.COEF_FLOAT32,
                dev.ptr(g.axes[0]), dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye),
                dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(pd), dev.ptr(iq), 0, 0, ns, nc, na, nt, c, fs, fd, fn,
            )
            ctx.save_for_backward(*packed)
            ctx.g = g
            ctx.c = c
            ctx.shape4 = (ns, nc, na, nt)
            ctx.c_dtype = sound_speed.dtype
            return iq.view(nx, nz, ny, nt)

        @staticmethod
        def backward(ctx, grad_iq):
            packed_rc, packed_cr = ctx.saved_tensors
            g = ctx.g
            ns, nc, na, nt = ctx.shape4
            _, fs, fd, fn = g.constants
            nx, nz, ny = g.shape
            grad = grad_iq.to(t.complex64).contiguous().view(g.nv, nt)
            g_packed = [t.zeros((nc, na, ns, nt), dtype=t.complex64, device=grad.device) for _ in range(2)]
            g_c = t.zeros((1,), dtype=t.float64, device=grad.device)
            lib = _lib.load()
            nbytes = int(lib.rcabeam_opw_adjoint_workspace_bytes(nx, nz, ny, ns, nc, na, nt))
            work = dev.empty(((nbytes + 7) // 8,), t.float64)
            _lib.call(
                "rcabeam_opw_adjoint_grid_deterministic_device", dev.ptr(packed_rc), dev.ptr(packed_cr),
                dev.ptr(g.axes[0]), dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye),
                dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(grad), dev.ptr(g_packed[0]), dev.ptr(g_packed[1]),
                dev.ptr(g_c), dev.ptr(work), nbytes, ns, nc, na, nt, ctx.c, fs, fd, fn,
            )
            out = []
            for gp in g_packed:
                gr = t.empty((ns, nc, na, nt), dtype=t.complex64, device=grad.device)
                _lib.call("rcabeam_unpack_ensemble_device", dev.ptr(gp), dev.ptr(gr), ns, nc, na, nt)
                out.append(gr)
            return out[0], out[1], g_c.reshape(()).to(ctx.c_dtype), None

    return OpwFunction


_OPW_FUNCTION = None


def opw_iq(rc, cr, angles, t_start, grid, geom: RCAGeometry, sound_speed=None):
    """Differentiable OPW IQ `(nx, nz, ny, frames)` complex64 on the GPU.

    rc, cr: complex64 device tensors `(samples, channels, angles, frames)`
    (requires_grad as needed).  sound_speed: optional 0-dim tensor overriding
    `geom.c` (requires_grad for dL/dc).  Up to 16 angles.
    """
    global _OPW_FUNCTION
    t = dev.torch()
    if _OPW_FUNCTION is None:
        _OPW_FUNCTION = _function()
    dev.require_default_stream("OPW autograd requires the default stream")
    if not (dev.is_tensor(rc) and dev.is_tensor(cr)) or rc.dtype != t.complex64 or cr.dtype != t.complex64:
        raise ValueError("rc and cr must be complex64 device tensors")
    if rc.shape != cr.shape or rc.ndim != 4 or rc.shape[0] < 2:
        raise ValueError("RC/CR must have matching nonempty (samples, channels, angles, frames) shapes")
    if rc.shape[2] > GRID_MAX_ANGLES:
        raise ValueError(f"At most {GRID_MAX_ANGLES} angles")
    if len(angles) != rc.shape[2] or len(geom.x_el) != rc.shape[1] or len(geom.y_el) != rc.shape[1]:
        raise ValueError("Geometry must match the channel/angle dimensions")
    _lib.load()
    g = _Geometry(np.asarray(angles), t_start, grid, geom)
    if sound_speed is None:
        sound_speed = t.tensor(float(geom.c), dtype=t.float64, device=rc.device)
    return _OPW_FUNCTION.apply(rc, cr, sound_speed, g)
