"""The full-size workload (128+128 elements, IQ 1024x128x11x200 per aperture, grid 128x256x128, the
GPU-simulated tissue + blood + noise mixture, seed 42) checked where the float64 oracle can follow:

* float32 grid OPW IQ and power at 600 seeded random voxels against the float64 oracle (rel L2 1e-5);
* the float16 low-rank form at the same voxels against the oracle (IQ within its 16-bit residual
  rounding: rel L2 < 1e-3; coded and uncoded forms differ only there);
* linearity at full size: OPW of rc + 2 cr equals OPW(rc, 0) + 2 OPW(0, cr) on the grid kernel;
* the clutter stage at full size: the GPU Gram, float64 EVD and filtered power of the whole float32 IQ
  (4,194,304 x 200) against the float64 oracle run on the same IQ, with the rank cut's eigen-gap.
"""

import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu

VOXELS = 600


@pytest.fixture(scope="module")
def full_scene():
    from rcabeam_amd.scene import mixture_scene

    print("case: full-size mixture scene 128x256x128, 200 frames", flush=True)
    scene, rc, cr = mixture_scene(grid_shape=(128, 256, 128))
    return scene, rc, cr


def _oracle_at(scene, rc, cr, pick):
    from oracle import das

    scan = das.scan_coords(scene.grid)
    g = scene.geom
    return das.opw_iq(rc.cpu().numpy(), cr.cpu().numpy(), scan[pick], g.x_el, g.y_el, scene.angles, scene.t_start,
                      g.c, g.fs, g.f_demod, g.fnumber)


def test_full_size_float32_grid_matches_oracle_at_sampled_voxels(full_scene):
    import torch

    import rcabeam_amd as rb

    scene, rc, cr = full_scene
    iq = rb.beamform_ensemble(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom)
    pd = rb.ensemble_pd_from_channels(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom)
    nt = iq.shape[-1]
    pick = np.sort(np.random.default_rng(1).choice(iq[..., 0].numel(), size=VOXELS, replace=False))
    expected = _oracle_at(scene, rc, cr, pick)
    got = iq.reshape(-1, nt)[torch.from_numpy(pick).cuda()].cpu().numpy()
    e_iq = rel_l2(got, expected)
    e_pd = rel_l2(pd.reshape(-1)[torch.from_numpy(pick).cuda()].cpu().numpy(), (np.abs(expected) ** 2).mean(-1))
    print(f"full size float32: IQ rel L2 {e_iq:.2e}, power {e_pd:.2e} at {VOXELS} voxels", flush=True)
    assert e_iq < 1e-5 and e_pd < 1e-5


def test_full_size_lowrank_form_matches_oracle_at_sampled_voxels(full_scene):
    import torch

    from rcabeam_amd.lowrank import beamform_lowrank

    scene, rc, cr = full_scene
    iq, pd = beamform_lowrank(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom, storage="float16", k=4)
    pick = np.sort(np.random.default_rng(2).choice(iq.shape[0], size=VOXELS, replace=False))
    expected = _oracle_at(scene, rc, cr, pick)
    index = torch.from_numpy(pick).cuda()
    e_iq = rel_l2(iq[index].cpu().numpy(), expected)
    e_pd = rel_l2(pd[index, 0].cpu().numpy(), (np.abs(expected) ** 2).mean(-1))
    print(f"full size float16 low-rank: IQ rel L2 {e_iq:.2e}, power {e_pd:.2e}", flush=True)
    assert e_iq < 1e-3 and e_pd < 1e-3


def test_full_size_grid_opw_is_linear(full_scene):
    import torch

    import rcabeam_amd as rb

    scene, rc, cr = full_scene
    zeros = torch.zeros_like(rc)
    args = (scene.angles, scene.t_start, scene.grid, scene.geom)
    both = rb.beamform_ensemble(rc, 2 * cr, *args)
    parts = rb.beamform_ensemble(rc, zeros, *args) + 2 * rb.beamform_ensemble(zeros, cr, *args)
    err = float(torch.linalg.vector_norm(both - parts) / torch.linalg.vector_norm(both))
    print(f"full size linearity: rel L2 {err:.2e}", flush=True)
    assert err < 2e-6  # measured 7.5e-7: float32 rounding of the two summation orders


def test_full_size_clutter_stage_matches_oracle(full_scene):
    """Exact float64 Gram + EVD + filtered power (rank_start 3) of the full-size float32 IQ against
    `oracle.clutter` on the same IQ (reference clutter.py:18-62, gpu_clutter.py:112-273).

    Bars: singular values rel 1e-9 (both sides form the Gram of the same complex64 values in float64;
    only summation order differs); the discarded subspace's projector within 1e-8; filtered power of
    every voxel rel L2 1e-5.  The rank cut is bit-exact when the retained index set is the same on
    both sides, which holds when the gap sigma_3 / sigma_4 (one-based) exceeds 1 by far more than the
    eigenvalue error: asserted with a factor 1e3.  Then, at sampled voxels, the GPU filtered power is
    compared with the oracle's filtered power of the oracle's float64 IQ; there the float32 IQ
    rounding e_iq passes through the filter unattenuated (it is white across frames), so the bound is
    2 e_iq sqrt(P_unfiltered / P_filtered) (first order), times 2 for margin.
    """
    import rcabeam_amd as rb
    from oracle import clutter as oclutter
    from rcabeam_amd import clutter

    scene, rc, cr = full_scene
    iq = rb.beamform_ensemble(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom)
    nt = iq.shape[-1]
    x = iq.reshape(-1, nt)
    values, basis = clutter.temporal_basis(x, precision="complex128")
    power = clutter.filtered_power(x, basis, rank_start=3).cpu().numpy()
    values, basis = values.cpu().numpy(), basis.cpu().numpy()
    print("case: full-size clutter stage, oracle Gram on the host", flush=True)
    host = x.cpu().numpy()
    del x
    ov, ob = oclutter.temporal_basis(host, method="gram")
    e_values = float(np.max(np.abs(values[:16] - ov[:16]) / ov[:16]))
    gap = float(ov[2] / ov[3])
    projector = basis[:, :3] @ basis[:, :3].conj().T
    o_projector = ob[:, :3] @ ob[:, :3].conj().T
    e_projector = float(np.linalg.norm(projector - o_projector))
    chunk = 1 << 18
    expected = np.concatenate([oclutter.filtered_power(host[i : i + chunk], ob, rank_start=3)
                               for i in range(0, len(host), chunk)])
    e_power = rel_l2(power, expected)
    print(f"full size clutter: sigma rel {e_values:.2e}, gap sigma3/sigma4 {gap:.4g} "
          f"(sigma1..5 {np.array2string(ov[:5], precision=4)}), projector {e_projector:.2e}, "
          f"filtered power rel L2 {e_power:.2e}", flush=True)
    assert e_values < 1e-9
    assert gap - 1 > 1e3 * max(e_values, 1e-12)
    assert e_projector < 1e-8
    assert e_power < 1e-5

    pick = np.sort(np.random.default_rng(3).choice(len(host), size=VOXELS, replace=False))
    exact_iq = _oracle_at(scene, rc, cr, pick)
    e_iq = rel_l2(host[pick], exact_iq)
    p_exact = oclutter.filtered_power(exact_iq, ob, rank_start=3)
    ratio = float(np.sum(np.abs(exact_iq) ** 2) / nt / np.sum(p_exact))
    e_sampled = rel_l2(power[pick], p_exact)
    bound = 4 * e_iq * np.sqrt(ratio)
    print(f"full size filtered power vs float64 IQ at {VOXELS} voxels: {e_sampled:.2e} (IQ {e_iq:.2e}, "
          f"P_unfiltered / P_filtered {ratio:.3g}, bound {bound:.2e})", flush=True)
    assert e_sampled < bound
