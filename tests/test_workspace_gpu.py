"""OpwWorkspace ownership, precision, traversal and context checks (reference tests/test_gpu_workspace.py)."""

import numpy as np
import pytest
from conftest import rel_l2

import rcabeam_amd as rb

pytestmark = pytest.mark.gpu


def _allocations():
    import torch

    return torch.cuda.memory_stats().get("allocation.all.allocated", 0)


@pytest.mark.parametrize("storage", ["float32", "float16", "bfloat16"])
@pytest.mark.parametrize("spatial", [False, True])
@pytest.mark.parametrize("frames", [3, 8, 66])
def test_workspace_reuses_buffers(storage, spatial, frames):
    """Reference tests/test_gpu_workspace.py:32-67: repeat reconstructions keep buffer identity and values."""
    import torch

    from rcabeam_amd.gpu import OpwWorkspace

    rng = np.random.default_rng(222)
    shape = (48, 3, 4, frames)
    rc = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    cr = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    elements = np.linspace(-0.001, 0.001, 3)
    geom = rb.RCAGeometry(elements, elements, fs=2e6, f_demod=1e6)
    angles = np.linspace(-0.1, 0.1, 4)
    grid = (np.linspace(-0.001, 0.001, 3), np.array([0.005, 0.007]), np.linspace(-0.001, 0.003, 5))
    expected = rb.beamform_ensemble(rc, cr, angles, 1e-6, grid, geom, iq_storage=storage)
    workspace = OpwWorkspace(shape, angles, 1e-6, grid, geom, iq_storage=storage, spatial_order=spatial)
    with pytest.raises(RuntimeError, match="Pack"):
        workspace.reconstruct()
    workspace.pack(torch.from_numpy(rc).cuda(), torch.from_numpy(cr).cuda())
    actual = workspace.reconstruct()
    assert actual.shape == (*tuple(len(a) for a in grid), frames) and actual.dtype == torch.complex64
    assert rel_l2(actual.cpu().numpy(), expected) < 1e-6
    power = workspace.power.cpu().numpy()
    assert rel_l2(power, np.mean(np.abs(expected) ** 2, axis=-1)) < 1e-6
    pointer = actual.data_ptr()
    angles[:] = 999  # metadata is a construction-time snapshot
    before = _allocations()
    assert workspace.reconstruct().data_ptr() == pointer
    assert _allocations() == before, "reconstruct allocated device memory"
    assert rel_l2(actual.cpu().numpy(), expected) < 1e-6
    workspace.pack(torch.from_numpy(rc * 2).cuda(), torch.from_numpy(cr * 2).cuda())
    assert workspace.reconstruct().data_ptr() == pointer
    assert rel_l2(actual.cpu().numpy(), expected * 2) < 1e-6


def test_workspace_rejects_stale_packs_and_streams():
    """Reference tests/test_gpu_workspace.py:70-92."""
    import torch

    from rcabeam_amd.gpu import OpwWorkspace

    shape = (48, 1, 1, 3)
    axis = np.array([0.0])
    geom = rb.RCAGeometry(axis, axis, fs=2e6, f_demod=1e6)
    grid = (axis, np.array([0.005]), axis)
    data = torch.zeros(shape, dtype=torch.complex64, device="cuda")
    workspace = OpwWorkspace(shape, axis, 0.0, grid, geom, iq_storage="float16")
    workspace.pack(data, data)
    workspace.reconstruct()
    data[0, 0, 0, 0] = 70000
    with pytest.raises(ValueError, match="FP16 IQ"):
        workspace.pack(data, data)
    with pytest.raises(RuntimeError, match="Pack"):
        workspace.reconstruct()
    with pytest.raises(RuntimeError, match="reading power"):
        _ = workspace.power
    with pytest.raises(ValueError, match="complex64 CUDA"):
        workspace.pack(np.zeros(shape, np.complex64), data)
    with torch.cuda.stream(torch.cuda.Stream()), pytest.raises(ValueError, match="default CUDA stream"):
        workspace.reconstruct()
    with pytest.raises(ValueError, match="IQ storage"):
        OpwWorkspace(shape, axis, 0.0, grid, geom, iq_storage="invalid")


def test_workspace_spatial_tile_boundary():
    """Reference tests/test_gpu_workspace.py:95-107: partial tiles and odd lateral tails."""
    import torch

    from rcabeam_amd.gpu import OpwWorkspace

    shape = (32, 2, 2, 8)
    rng = np.random.default_rng(22)
    rc = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    el = np.array([-0.001, 0.001])
    angles = np.array([-0.1, 0.1])
    grid = (np.linspace(-0.001, 0.001, 65), np.array([0.005]), np.linspace(-0.001, 0.001, 253))
    geom = rb.RCAGeometry(el, el, fs=2e6, f_demod=1e6)
    workspace = OpwWorkspace(shape, angles, 0.0, grid, geom, spatial_order=True)
    data = torch.from_numpy(rc).cuda()
    workspace.pack(data, data)
    expected = rb.beamform_ensemble(rc, rc, angles, 0.0, grid, geom)
    assert rel_l2(workspace.reconstruct().cpu().numpy(), expected) < 1e-6
