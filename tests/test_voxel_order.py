"""`opw_voxel_order` on the CPU: the cases of reference tests/test_opw_order.py:9-20.

The permutation is integer index work, so every check is exact.
"""

import numpy as np
import pytest

from rcabeam_amd.ensemble import opw_voxel_order


@pytest.mark.parametrize("shape", [(4, 3, 8), (3, 2, 5), (1, 2, 1), (2, 1, 4), (5, 3, 2)])
def test_order_is_a_permutation_that_scatters_back(shape):
    order = opw_voxel_order(shape)
    expected = np.arange(np.prod(shape))
    np.testing.assert_array_equal(np.sort(order), expected)
    restored = np.empty_like(order)
    restored[order] = expected[order]
    np.testing.assert_array_equal(restored, expected)


def test_first_patch_known_answer():
    """The first 2 (x) by 4 (y) patch of a (4, 3, 8) grid: x = 0 then x = 1 (voxel id + nz * ny = 24)."""
    np.testing.assert_array_equal(opw_voxel_order((4, 3, 8))[:8], [0, 1, 2, 3, 24, 25, 26, 27])


def test_patches_then_tails_by_brute_force():
    """Full order of a (5, 2, 6) grid written out from the definition: whole 2 x 4 patches in (x pair,
    z, y quad) order, then the y tail of the paired x lines, then the odd last x line."""
    nx, nz, ny = 5, 2, 6
    vid = lambda ix, iz, iy: (ix * nz + iz) * ny + iy  # noqa: E731
    expected = []
    for px in range(nx // 2):
        for iz in range(nz):
            for py in range(ny // 4):
                for dx in range(2):
                    expected += [vid(2 * px + dx, iz, 4 * py + dy) for dy in range(4)]
    expected += [vid(ix, iz, iy) for ix in range(nx - nx % 2) for iz in range(nz) for iy in range(ny - ny % 4, ny)]
    expected += [vid(ix, iz, iy) for ix in range(nx - nx % 2, nx) for iz in range(nz) for iy in range(ny)]
    np.testing.assert_array_equal(opw_voxel_order((nx, nz, ny)), expected)


def test_rejects_empty_grid():
    with pytest.raises(ValueError):
        opw_voxel_order((0, 2, 4))
