"""CPU checks that pin the oracle's distance field (orc_voxelize,
orc_edt_brute) before it judges the GPU transform: scipy's exact Euclidean
distance transform, closed forms, the voxel rounding rule, and the
reference's own test on its fixture (geometry/tests/cuda_distance_field_test.cpp)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def fixture_points():
    """tests/golden/random_grid.voxels.txt (the reference's
    test_assets/random_grid.voxels.txt): (3, 216) points."""
    return np.loadtxt(os.path.join(GOLDEN, "random_grid.voxels.txt"), dtype=np.float32).T


REF_ORIGIN = np.array([-0.75, -0.75, 0.0], np.float32)
REF_VOXEL, REF_G = np.float32(0.02), 128


@pytest.mark.parametrize("G,density,seed", [(24, 0.002, 0), (24, 0.05, 1), (31, 0.3, 2),
                                            (17, 0.9, 3)])
def test_edt_matches_scipy(orc, G, density, seed):
    from scipy.ndimage import distance_transform_edt
    rng = np.random.default_rng(seed)
    occ = (rng.random((G, G, G)) < density).astype(np.uint8)
    occ[rng.integers(G), rng.integers(G), rng.integers(G)] = 1
    d2 = orc.edt_brute(occ)
    ref = distance_transform_edt(1 - occ)
    np.testing.assert_array_equal(d2, np.rint(ref ** 2).astype(np.int64))


def test_edt_closed_forms(orc):
    G = 9
    occ = np.zeros((G, G, G), np.uint8)
    assert (orc.edt_brute(occ) == orc.FAR).all()  # nothing occupied
    occ[2, 7, 3] = 1  # z, y, x
    z, y, x = np.meshgrid(np.arange(G), np.arange(G), np.arange(G), indexing="ij")
    one = (x - 3) ** 2 + (y - 7) ** 2 + (z - 2) ** 2
    np.testing.assert_array_equal(orc.edt_brute(occ), one)
    occ[8, 0, 8] = 1
    two = np.minimum(one, (x - 8) ** 2 + y ** 2 + (z - 8) ** 2)
    np.testing.assert_array_equal(orc.edt_brute(occ), two)


def test_voxelize_rounding(orc):
    o = np.zeros(3, np.float32)
    vs = np.float32(0.25)
    pts = np.array([[0.125, 0.0, 0.0],     # exactly half a voxel: rounds away from zero
                    [-0.125, 0.0, 0.0],    # -0.5 -> -1: outside, ignored
                    [0.374, 0.5, 0.75],    # (1, 2, 3)
                    [9.0, 0.0, 0.0]], np.float32).T
    occ = orc.voxelize(pts, o, vs, 4)
    assert occ.sum() == 2 and occ[0, 0, 1] == 1 and occ[3, 2, 1] == 1


def test_reference_fixture(orc):
    """cuda_distance_field_test.cpp:16-79: every 4th voxel's distance to the
    nearest input point, within 1e-3 of a scan over all points."""
    pts = fixture_points()
    assert pts.shape == (3, 216)
    occ = orc.voxelize(pts, REF_ORIGIN, REF_VOXEL, REF_G)
    d = np.sqrt(orc.edt_brute(occ).astype(np.float32)) * REF_VOXEL
    idx = np.arange(0, REF_G, 4)
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    P = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32) * REF_VOXEL + REF_ORIGIN
    brute = np.sqrt(((P[:, None, :] - pts.T[None, :, :]) ** 2).sum(-1)).min(1)
    np.testing.assert_allclose(d[z, y, x].reshape(-1), brute, atol=1e-3)
