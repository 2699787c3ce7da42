"""Dense float64 torch restatement of the OPW delay-and-sum, differentiable in the
channel IQ and the sound speed: the autograd reference for the adjoint kernel.

Same semantics as `oracle/das.py` (reference src/rcabeam/sim.py:344-393) on the
float32-rounded inputs that cross the C ABI: receive delay sqrt(z^2 + (v - e)^2) / c,
transmit delay (z cos a + u sin a) / c, s = (tau - t0) fs valid on the closed window
[0, ns - 1], linear interpolation between floor(s) and min(floor(s) + 1, ns - 1),
rotation by exp(+2j pi fd tau), aperture mask |v - e| <= z / (2 F).  Test
infrastructure only: small grids.
"""

from __future__ import annotations

import numpy as np
import torch


def _f32(x):
    return torch.as_tensor(np.asarray(x, dtype=np.float32).astype(np.float64))


def opw_iq(rc, cr, scan, x_el, y_el, angles, starts, c, fs, fd, fnumber):
    """rc, cr: complex128 (ns, nc, na, nt) tensors; c: float64 0-dim tensor. Returns (nv, nt) complex128."""
    ns, nc, na, nt = rc.shape
    coords = _f32(scan)
    x, z, y = coords[:, 0], coords[:, 1], coords[:, 2]
    th = _f32(angles)
    t0 = torch.broadcast_to(_f32(starts), (na,))
    fs = float(np.float32(fs))
    fd = float(np.float32(fd))
    full = fnumber is None or float(np.float32(fnumber)) <= 0
    out = torch.zeros((len(coords), nt), dtype=torch.complex128)
    channels = torch.arange(nc)
    for data, el, rx_coord, tx_coord in ((rc, _f32(x_el), x, y), (cr, _f32(y_el), y, x)):
        distance = rx_coord[:, None] - el[None, :]
        receive = torch.sqrt(z[:, None] ** 2 + distance**2) / c
        active = torch.ones_like(distance, dtype=torch.bool) if full else distance.abs() <= z[:, None] / (
            2 * float(np.float32(fnumber)))
        for m in range(na):
            transmit = (z * torch.cos(th[m]) + tx_coord * torch.sin(th[m])) / c
            tau = transmit[:, None] + receive
            sample = (tau - t0[m]) * fs
            valid = (sample >= 0) & (sample <= ns - 1)
            clipped = sample.clamp(0, ns - 1)
            lower = torch.floor(clipped.detach()).long()
            upper = torch.clamp(lower + 1, max=ns - 1)
            w = (clipped - lower.to(torch.float64))[..., None]
            plane = data[:, :, m, :]
            a = plane[lower, channels[None, :]]
            b = plane[upper, channels[None, :]]
            value = a + w * (b - a)
            phase = torch.exp(1j * 2 * np.pi * fd * tau)
            keep = (valid & active).to(torch.float64)
            out = out + (value * (phase * keep)[..., None]).sum(dim=1)
    return out
