"""Markers: `gpu` tests need an MI355X (they call the HIP library through the C ABI)."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built rcabeam_amd library")


def pytest_collection_modifyitems(config, items):
    try:
        import torch

        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no HIP GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def rel_l2(actual, expected) -> float:
    """Relative L2 error ||a - e|| / ||e|| in float64 (0 when both are zero)."""
    import numpy as np

    a = np.asarray(actual).astype(np.complex128 if np.iscomplexobj(actual) else np.float64)
    e = np.asarray(expected).astype(np.complex128 if np.iscomplexobj(expected) else np.float64)
    denom = np.linalg.norm(e.ravel())
    num = np.linalg.norm((a - e).ravel())
    if denom == 0:
        return float(num)
    return float(num / denom)
