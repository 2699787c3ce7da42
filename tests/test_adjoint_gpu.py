"""Adjoint (transpose) OPW kernel behind torch.autograd: dot-product identity, float64
autograd reference, central finite differences.

Workload: 64 + 64 elements at 200 um, c = 1540 m/s, fs = 20 MHz, fd = 15 MHz, 3 angles,
256 samples, 32 frames (the point scene); gradient checks on an 8 x 16 x 8 sub-grid.
Loss: L2 distance between the clutter-filtered power Doppler (Gram + eigh in complex128,
rank_start 1) and a seeded target map, plus 0.1 x a sharpness term (negative normalized
4th moment of the power).
"""

import numpy as np
import pytest
import torch
import torch_reference as tref
from conftest import rel_l2

import rcabeam_amd as rb
from oracle import das, scene

pytestmark = pytest.mark.gpu


def _case(grid_shape=(8, 16, 8), frames=32, seed=5):
    acq = scene.point_scene(grid_shape=grid_shape, frames=frames, seed=seed)
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    return acq, geom


def _loss(iq, target, rank_start=1):
    nt = iq.shape[-1]
    x = iq.reshape(-1, nt).to(torch.complex128)
    gram = x.conj().T @ x
    _, vectors = torch.linalg.eigh(gram)
    keep = vectors[:, : nt - rank_start]  # ascending eigenvalues: drop the rank_start largest
    filtered = (x @ keep) @ keep.conj().T
    pd = (filtered.abs() ** 2).mean(dim=-1).reshape(target.shape)
    l2 = ((pd - target) ** 2).sum() / (target**2).sum()
    sharp = -(pd**4).sum() / (pd**2).sum() ** 2
    return l2 + 0.1 * sharp


def _target(shape, scale, seed=9):
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.uniform(0.5, 1.5, size=shape) * scale)


def test_adjoint_dot_product_identity():
    """<A x, y> = <x, A^H y> for random channels x and IQ gradients y (float32 rounding only)."""
    from rcabeam_amd.autograd import opw_iq

    acq, geom = _case(grid_shape=(12, 20, 10), frames=40)
    rng = np.random.default_rng(3)
    shape = acq.rc.shape
    xr = torch.from_numpy((rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)).cuda()
    xc = torch.from_numpy((rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)).cuda()
    xr.requires_grad_(True)
    xc.requires_grad_(True)
    iq = opw_iq(xr, xc, acq.angles, acq.t_start, acq.grid, geom)
    y = torch.from_numpy((rng.normal(size=iq.shape) + 1j * rng.normal(size=iq.shape)).astype(np.complex64)).cuda()
    gr, gc = torch.autograd.grad(iq, (xr, xc), grad_outputs=y)
    lhs = torch.vdot(y.reshape(-1).to(torch.complex128), iq.detach().reshape(-1).to(torch.complex128)).real
    rhs = (torch.vdot(gr.reshape(-1).to(torch.complex128), xr.detach().reshape(-1).to(torch.complex128))
           + torch.vdot(gc.reshape(-1).to(torch.complex128), xc.detach().reshape(-1).to(torch.complex128))).real
    print(f"<Ax,y> = {lhs.item():.9e}, <x,A^H y> = {rhs.item():.9e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_gradients_match_float64_autograd():
    """Channel and sound-speed gradients of the filtered-PD loss against dense float64 torch autograd."""
    from rcabeam_amd.autograd import opw_iq

    acq, geom = _case()
    shape = tuple(len(a) for a in acq.grid)
    rc = torch.from_numpy(acq.rc).cuda().requires_grad_(True)
    cr = torch.from_numpy(acq.cr).cuda().requires_grad_(True)
    c = torch.tensor(float(np.float32(acq.c)), dtype=torch.float64, device="cuda", requires_grad=True)
    iq = opw_iq(rc, cr, acq.angles, acq.t_start, acq.grid, geom, sound_speed=c)
    scale = float((iq.detach().abs() ** 2).mean())
    target = _target(shape, scale).cuda()
    loss = _loss(iq, target)
    loss.backward()

    rc64 = torch.from_numpy(acq.rc.astype(np.complex128)).requires_grad_(True)
    cr64 = torch.from_numpy(acq.cr.astype(np.complex128)).requires_grad_(True)
    c64 = torch.tensor(float(np.float32(acq.c)), dtype=torch.float64, requires_grad=True)
    scan = das.scan_coords(acq.grid)
    iq64 = tref.opw_iq(rc64, cr64, scan, acq.x_el, acq.y_el, acq.angles, acq.t_start, c64, acq.fs, acq.fd,
                       acq.fnumber).reshape(*shape, -1)
    assert rel_l2(iq.detach().cpu().numpy(), iq64.detach().numpy()) < 1e-5
    loss64 = _loss(iq64, target.cpu())
    loss64.backward()
    e_rc = rel_l2(rc.grad.cpu().numpy(), rc64.grad.numpy())
    e_cr = rel_l2(cr.grad.cpu().numpy(), cr64.grad.numpy())
    e_c = abs(c.grad.item() - c64.grad.item()) / abs(c64.grad.item())
    print(f"loss {loss.item():.6e} vs {loss64.item():.6e}; grad rel err rc {e_rc:.2e} cr {e_cr:.2e} "
          f"c {e_c:.2e} (dL/dc {c.grad.item():.6e} vs {c64.grad.item():.6e})", flush=True)
    assert e_rc < 1e-4 and e_cr < 1e-4
    assert e_c < 1e-4


def test_central_finite_differences():
    """Kernel gradients against central differences of the float64 reference loss.

    Channels: random directions, relative step 1e-3 (the loss is smooth in the channel IQ).
    Sound speed: relative step 1e-3 (the stated check) and 1e-7.  Linear interpolation is
    piecewise linear in c: a 1e-3 step moves every tap by up to ~0.3 samples, across sample
    boundaries where the interpolation slope changes, so that quotient is the slope averaged
    over the step and differs from the local derivative by a few tens of percent; at 1e-7 hardly
    any tap crosses a boundary and the quotient must match the local derivative."""
    from rcabeam_amd.autograd import opw_iq

    acq, geom = _case()
    shape = tuple(len(a) for a in acq.grid)
    rc = torch.from_numpy(acq.rc).cuda().requires_grad_(True)
    cr = torch.from_numpy(acq.cr).cuda().requires_grad_(True)
    c0 = float(np.float32(acq.c))
    c = torch.tensor(c0, dtype=torch.float64, device="cuda", requires_grad=True)
    iq = opw_iq(rc, cr, acq.angles, acq.t_start, acq.grid, geom, sound_speed=c)
    target = _target(shape, float((iq.detach().abs() ** 2).mean()))
    _loss(iq, target.cuda()).backward()
    scan = das.scan_coords(acq.grid)
    rc64 = torch.from_numpy(acq.rc.astype(np.complex128))
    cr64 = torch.from_numpy(acq.cr.astype(np.complex128))

    def loss64(a, b, speed):
        with torch.no_grad():
            iq64 = tref.opw_iq(a, b, scan, acq.x_el, acq.y_el, acq.angles, acq.t_start,
                               torch.tensor(speed, dtype=torch.float64), acq.fs, acq.fd, acq.fnumber)
            return float(_loss(iq64.reshape(*shape, -1), target))

    rng = np.random.default_rng(11)
    shape4 = acq.rc.shape
    dr = torch.from_numpy(rng.normal(size=shape4) + 1j * rng.normal(size=shape4)) * float(rc64.abs().mean())
    dc = torch.from_numpy(rng.normal(size=shape4) + 1j * rng.normal(size=shape4)) * float(cr64.abs().mean())
    eps = 1e-3
    fd_dir = (loss64(rc64 + eps * dr, cr64 + eps * dc, c0) - loss64(rc64 - eps * dr, cr64 - eps * dc, c0)) / (2 * eps)
    analytic = float((torch.vdot(rc.grad.cpu().to(torch.complex128).reshape(-1), dr.reshape(-1)).real
                      + torch.vdot(cr.grad.cpu().to(torch.complex128).reshape(-1), dc.reshape(-1)).real))
    dldc = c.grad.item()
    fd_c = {}
    for rel in (1e-3, 1e-7):
        step = rel * c0
        fd_c[rel] = (loss64(rc64, cr64, c0 + step) - loss64(rc64, cr64, c0 - step)) / (2 * step)
    print(f"channels: fd {fd_dir:.8e} analytic {analytic:.8e}; dL/dc analytic {dldc:.6e}, "
          f"fd(1e-3) {fd_c[1e-3]:.6e}, fd(1e-7) {fd_c[1e-7]:.6e}", flush=True)
    assert abs(fd_dir - analytic) <= 1e-4 * abs(analytic)
    assert abs(fd_c[1e-7] - dldc) <= 1e-3 * abs(dldc)
    assert np.sign(fd_c[1e-3]) == np.sign(dldc) and abs(fd_c[1e-3] - dldc) <= 0.5 * abs(dldc)


def test_adjoint_is_deterministic_and_matches_the_atomic_form():
    """The autograd backward (fixed-point accumulation, fixed-order dL/dc) gives bit-identical gradients on
    repeated calls, and agrees with the float-atomic adjoint entry point to float32 summation order
    (rel L2 1e-6)."""
    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.autograd import opw_iq
    from rcabeam_amd.ensemble import _Geometry, pack_channels

    acq, geom = _case(grid_shape=(16, 24, 16), frames=40)
    rng = np.random.default_rng(8)
    xr = torch.from_numpy(acq.rc).cuda().requires_grad_(True)
    xc = torch.from_numpy(acq.cr).cuda().requires_grad_(True)
    c = torch.tensor(1540.0, dtype=torch.float64, device="cuda", requires_grad=True)
    iq = opw_iq(xr, xc, acq.angles, acq.t_start, acq.grid, geom, sound_speed=c)
    y = torch.from_numpy((rng.normal(size=iq.shape) + 1j * rng.normal(size=iq.shape)).astype(np.complex64)).cuda()
    runs = [torch.autograd.grad(iq, (xr, xc, c), grad_outputs=y, retain_graph=True) for _ in range(3)]
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert torch.equal(a, b)
    # The float-atomic entry point on the same packed channels (accumulation order differs run to run).
    g = _Geometry(acq.angles, acq.t_start, acq.grid, geom)
    ns, nc, na, nt = acq.rc.shape
    packed = [pack_channels(x.detach(), "float32") for x in (xr, xc)]
    grads = [torch.zeros((nc, na, ns, nt), dtype=torch.complex64, device="cuda") for _ in range(2)]
    gc = torch.zeros((1,), dtype=torch.float64, device="cuda")
    nx, nz, ny = g.shape
    _, fs, fd, fn = g.constants
    _lib.call("rcabeam_opw_adjoint_grid_device", dev.ptr(packed[0]), dev.ptr(packed[1]), dev.ptr(g.axes[0]),
              dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles),
              dev.ptr(g.starts), dev.ptr(y.reshape(g.nv, nt).contiguous()), dev.ptr(grads[0]), dev.ptr(grads[1]),
              dev.ptr(gc), ns, nc, na, nt, 1540.0, fs, fd, fn)
    unpacked = grads[0].permute(2, 0, 1, 3).contiguous()  # (nc, na, ns, nt) -> (ns, nc, na, nt)
    assert rel_l2(runs[0][0].cpu().numpy(), unpacked.cpu().numpy()) < 1e-6
    assert abs(float(runs[0][2]) - float(gc[0])) <= 1e-6 * abs(float(gc[0]))
