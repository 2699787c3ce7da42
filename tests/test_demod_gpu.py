"""`rcabeam_rf_to_iq_device` against the float64 demodulation oracle, and RF -> IQ -> OPW end to end.

Bars: rel L2 1e-5 for float32 and int16 RF (float32 accumulation of up to 63 complex taps of float32
products); host and packed layouts, edge rows included.  End to end, an RF point echo demodulated on the
GPU and beamformed equals the beamformed IQ of the same echo demodulated by the oracle (1e-5).
"""

import numpy as np
import pytest
from conftest import rel_l2

from oracle import demod

pytestmark = pytest.mark.gpu

FS, FD = 60e6, 15e6


@pytest.mark.parametrize("rf_type", ["float32", "int16"])
@pytest.mark.parametrize("ntaps,dec", [(31, 3), (63, 3), (1, 1), (15, 4), (63, 1)])
def test_rf_to_iq_matches_oracle(rf_type, ntaps, dec):
    """Polyphase kernel (taps per phase <= 32) and the direct kernel ((63, 1)); 273 columns (two column
    blocks, the second partial) and 301 RF samples, so most output blocks touch the record's ends."""
    import torch

    from rcabeam_amd.demod import lowpass_taps, rf_to_iq

    rng = np.random.default_rng(ntaps + dec)
    shape = (301, 13, 3, 7)
    rf = rng.normal(size=shape) * (1000 if rf_type == "int16" else 1)
    rf = rf.astype(np.int16 if rf_type == "int16" else np.float32)
    taps = lowpass_taps(ntaps, FS / (2 * dec) * 0.9, FS) if ntaps > 1 else np.ones(1, np.float32)
    t0 = 2.7e-6
    print(f"case: rf_to_iq {rf_type} ntaps={ntaps} dec={dec}", flush=True)
    expected = demod.rf_to_iq(rf, taps, dec, FS, FD, t0)
    got = rf_to_iq(rf, FS, FD, dec, taps=taps, t_start=t0)
    assert got.shape == expected.shape and got.dtype == np.complex64
    assert rel_l2(got, expected) < 1e-5
    packed = rf_to_iq(torch.from_numpy(rf).cuda(), FS, FD, dec, taps=taps, t_start=t0, packed=True)
    assert tuple(packed.shape) == (13, 3, expected.shape[0], 7)
    assert rel_l2(packed.cpu().numpy(), expected.transpose(1, 2, 0, 3)) < 1e-5


def test_rf_to_iq_rejects_bad_arguments():
    from rcabeam_amd.demod import rf_to_iq

    rf = np.zeros((30, 1, 1, 2), np.float32)
    with pytest.raises(ValueError):
        rf_to_iq(rf, FS, FD, 3, taps=np.ones(4, np.float32))  # even tap count
    with pytest.raises(ValueError):
        rf_to_iq(rf.astype(np.float64), FS, FD, 3)
    with pytest.raises(ValueError):
        rf_to_iq(rf, FS, FD, 3, ns_iq=11)  # beyond ceil(30 / 3)


def test_rf_point_echo_through_demodulation_and_opw():
    """RF of a point scatterer (Gaussian envelope on the 15 MHz carrier, 64 + 64 elements, 3 angles) ->
    GPU demodulation into packed channels -> grid OPW, against the oracle's demodulation of the same RF
    beamformed by the float64 delay-and-sum oracle."""
    import torch

    import rcabeam_amd as rb
    from oracle import das, scene

    acq = scene.point_scene(grid_shape=(8, 16, 8), frames=4)
    el, angles = acq.x_el, acq.angles
    ns_rf, dec = 768, 3
    t = np.arange(ns_rf) / FS
    rf = []
    rng = np.random.default_rng(3)
    point = (0.4e-3, 7.1e-3, -0.6e-3)
    phases = np.exp(1j * rng.uniform(-np.pi, np.pi, 4))
    for config in ("RC", "CR"):
        v = point[0] if config == "RC" else point[2]
        u = point[2] if config == "RC" else point[0]
        out = np.zeros((ns_rf, len(el), len(angles), 4), np.float32)
        for a, th in enumerate(angles):
            tau = (point[1] * np.cos(th) + u * np.sin(th)) / acq.c + np.sqrt(point[1] ** 2 + (v - el) ** 2) / acq.c
            env = np.exp(-(((t[:, None] - tau[None, :]) / 0.15e-6) ** 2))
            carrier = np.exp(1j * 2 * np.pi * FD * (t[:, None] - tau[None, :]))
            sig = env[..., None] * carrier[..., None] * phases[None, None, :]
            out[:, :, a, :] = sig.real
        rf.append(out)
    from rcabeam_amd.demod import lowpass_taps, rf_to_iq

    taps = lowpass_taps(63, 8e6, FS)
    iq_ref = [demod.rf_to_iq(x, taps, dec, FS, FD).astype(np.complex64) for x in rf]
    geom = rb.RCAGeometry(el, el, c=acq.c, fs=FS / dec, f_demod=FD, fnumber=1.0)
    iq_gpu = [rf_to_iq(torch.from_numpy(x).cuda(), FS, FD, dec, taps=taps) for x in rf]
    assert rel_l2(iq_gpu[0].cpu().numpy(), iq_ref[0]) < 1e-5
    beam = rb.beamform_ensemble(iq_gpu[0], iq_gpu[1], angles, 0.0, acq.grid, geom).cpu().numpy()
    expected = das.opw_iq(iq_ref[0], iq_ref[1], das.scan_coords(acq.grid), el, el, angles, 0.0, acq.c, FS / dec, FD,
                          1.0)
    assert rel_l2(beam.reshape(expected.shape), expected) < 1e-5
