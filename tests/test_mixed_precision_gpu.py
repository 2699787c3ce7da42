"""Mixed-precision OPW + SVD clutter filter + power Doppler against the float64 oracle.

Scene: the mixture of the full-size benchmark at reduced size (same generator, seed 42):
64 + 64 elements at 200 um, 640 samples, 5 angles in +/-10 deg, 64 frames, 4000 static
Rayleigh tissue scatterers (rank-3 temporal clutter with a slow common phase drift), 400 blood
scatterers at -40 dB moving at 1-20 mm/s in three tubes at PRF 500 Hz, -60 dB noise; grid
8 x 24 x 8 over +/-6.4 mm x 4-20 mm; clutter rank_start 3.

Every configuration runs the GPU path end to end (pack -> grid delay-and-sum -> Gram -> float64
EVD -> filtered power) and is compared with the float64 oracle (delay-and-sum, SVD basis,
filtered power) on the same complex64 channels.  Relative L2 errors of IQ, unfiltered power
Doppler and filtered power and relative errors of the three clutter singular values are printed
and bounded (bounds sit just above the measured errors, as regression guards).

16-bit channel storage quantizes the 0 dB tissue echoes with a relative error of ~2^-12 (FP16)
or ~2^-9 (BF16).  That error is white across frames, so the clutter filter cannot remove it and it
lands on the -40 dB blood signal: filtered power errors of a few 1e-3 (FP16) and several 1e-2
(BF16), while IQ and unfiltered power stay below 1e-3 with FP16.  Low-rank + residual coding
(rcabeam_amd.lowrank, k = 4 temporal modes of the channels kept as 16-bit hi + lo coefficients)
rounds only the residual: filtered power 4e-5 (FP16) and 3e-4 (BF16) with complex64 IQ.  The
beamformed IQ stored in FP16 rounds the tissue again, so that form stays at a few 1e-3.
"""

import numpy as np
import pytest
import torch
from conftest import rel_l2

from oracle import clutter as oc
from oracle import das

pytestmark = pytest.mark.gpu

RANK_START = 3
_CACHE = {}


def _scene():
    if "scene" not in _CACHE:
        from rcabeam_amd.scene import mixture_scene

        scene, rc, cr = mixture_scene(elements=64, samples=640, angle_count=5, frames=64, grid_shape=(8, 24, 8),
                                      lateral=6.4e-3, n_tissue=4000, n_blood=400)
        g = scene.geom
        scan = das.scan_coords(scene.grid)
        rc_h, cr_h = rc.cpu().numpy(), cr.cpu().numpy()
        iq64 = das.opw_iq(rc_h, cr_h, scan, g.x_el, g.y_el, scene.angles, scene.t_start, g.c, g.fs, g.f_demod,
                          g.fnumber)
        values64, basis64 = oc.temporal_basis(iq64, method="svd")
        power64 = oc.filtered_power(iq64, basis64, rank_start=RANK_START)
        _CACHE["scene"] = (scene, rc, cr, iq64, values64, power64)
    return _CACHE["scene"]


def _run(channel_storage, iq_storage, coefficients, gram_precision, lowrank_k=0):
    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd import clutter as gc
    from rcabeam_amd.ensemble import _Geometry, pack_channels
    from rcabeam_amd.lowrank import beamform_lowrank

    scene, rc, cr, *_ = _scene()
    lib = _lib.load()
    g = _Geometry(scene.angles, scene.t_start, scene.grid, scene.geom)
    ns, nc, na, nt = rc.shape
    code = _lib.STORAGE[iq_storage]
    if lowrank_k:
        signal, pd = beamform_lowrank(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom,
                                      storage=channel_storage, k=lowrank_k, coefficients=coefficients,
                                      iq_storage=iq_storage)
    else:
        packed = [pack_channels(x, channel_storage) for x in (rc, cr)]
        pd = dev.empty((g.nv, 2), torch.float32)
        signal = dev.empty((g.nv, nt), torch.complex64 if code == 0 else torch.int32)
        nx, nz, ny = g.shape
        c, fs, fd, fn = g.constants
        _lib.check(lib.rcabeam_opw_grid_device(
            dev.ptr(packed[0]), dev.ptr(packed[1]), _lib.STORAGE[channel_storage], coefficients,
            dev.ptr(g.axes[0]), dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe), dev.ptr(g.ye),
            dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(pd), dev.ptr(signal), code, 0, ns, nc, na, nt,
            c, fs, fd, fn), "das")
    storage = None if code == 0 else iq_storage
    values, basis = gc.temporal_basis(signal, precision=gram_precision, storage=storage)
    power = gc.filtered_power(signal, basis, rank_start=RANK_START, storage=storage)
    if code == 0:
        iq = signal
    else:
        half = torch.float16 if iq_storage == "float16" else torch.bfloat16
        iq = signal.view(half).to(torch.float32).view(torch.complex64)
    return iq.cpu().numpy(), np.asarray(values.cpu() if torch.is_tensor(values) else values), \
        np.asarray(power.cpu() if torch.is_tensor(power) else power), pd[:, 0].cpu().numpy()


# channels, beamformed IQ, coefficients (0: float32-accurate, 1: one 16-bit per sample), Gram precision,
# bounds on (IQ, unfiltered power, filtered power, clutter singular values), low-rank coding modes k
# (0: channels rounded directly)
CASES = {
    "float32": ("float32", "float32", 0, "complex128", (1e-5, 1e-5, 2e-5, 1e-6), 0),
    "fp16 channels + fp16 IQ": ("float16", "float16", 1, "complex64", (5e-4, 1e-3, 6e-3, 1e-5), 0),
    "fp16 channels + complex64 IQ": ("float16", "float32", 1, "complex64", (5e-4, 1e-3, 4e-3, 1e-5), 0),
    "bf16 channels + bf16 IQ": ("bfloat16", "bfloat16", 1, "complex64", (4e-3, 8e-3, 2e-1, 1e-4), 0),
    "bf16 channels + complex64 IQ": ("bfloat16", "float32", 1, "complex64", (3e-3, 6e-3, 5e-2, 1e-4), 0),
    "fp16 low-rank channels + complex64 IQ": ("float16", "float32", 1, "complex128", (4e-4, 3e-4, 1.5e-4, 5e-6), 4),
    "fp16 low-rank channels + complex64 IQ, split Gram": ("float16", "float32", 1, "complex64",
                                                          (4e-4, 3e-4, 1.5e-4, 1e-5), 4),
    "fp16 low-rank channels + fp16 IQ": ("float16", "float16", 1, "complex64", (6e-4, 3e-4, 8e-3, 1e-5), 4),
    "bf16 low-rank channels + complex64 IQ": ("bfloat16", "float32", 1, "complex128", (3e-3, 2e-3, 1e-3, 2e-4), 4),
    "fp16 low-rank channels, float32-accurate coefficients": ("float16", "float32", 0, "complex128",
                                                              (1e-5, 1e-5, 5e-5, 5e-7), 4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_pipeline_error_against_float64_oracle(name):
    from oracle import combiners

    channel_storage, iq_storage, coefficients, precision, bounds, lowrank_k = CASES[name]
    _, _, _, iq64, values64, power64 = _scene()
    iq, values, power, pd = _run(channel_storage, iq_storage, coefficients, precision, lowrank_k)
    e_iq = rel_l2(iq, iq64)
    e_pd = rel_l2(pd, combiners.power_doppler(iq64))
    e_fp = rel_l2(power, power64)
    e_sv = np.max(np.abs(values[:RANK_START] - values64[:RANK_START]) / values64[:RANK_START])
    print(f"{name}: IQ {e_iq:.2e}, power {e_pd:.2e}, filtered power {e_fp:.2e}, "
          f"clutter singular values {e_sv:.2e}", flush=True)
    assert e_iq < bounds[0] and e_pd < bounds[1] and e_fp < bounds[2] and e_sv < bounds[3]
