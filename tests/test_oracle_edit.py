"""CPU checks that pin the oracle's region editing (EditRegionsCuda,
planning/cuda_edit_regions.cu:44-300): hand-derived faces, ordering and
redundancy cases, and the reference's own test scenario
(planning/tests/cuda_edit_regions_test.cpp:74-216) with its assertions."""
import numpy as np
import pytest

from conftest import parse

SEG = (np.array([-1.0, 0.0], np.float32), np.array([1.0, 0.0], np.float32))


def _box(lo, hi):
    """MakeBox rows, interleaved (-e_i, +e_i) (hpolyhedron.cpp:104-122)."""
    d = len(lo)
    A = np.zeros((2 * d, d), np.float32)
    b = np.zeros(2 * d, np.float32)
    for i in range(d):
        A[2 * i, i], A[2 * i + 1, i] = -1, 1
        b[2 * i], b[2 * i + 1] = -lo[i], hi[i]
    return A, b


def _in_set(A, b, x, tol=1e-9):
    """PointInSet (hpolyhedron.cpp:265-274) in float."""
    x = np.asarray(x, np.float32)
    return bool(np.all(A.astype(np.float32) @ x <= b + np.float32(tol)))


def test_face_closed_form(orc):
    A, b = _box([-2, -2], [2, 2])
    # collision straight above its projection: normal +y, offset 1.0 - margin
    A2, b2, added = orc.edit_region(A, b, [[0.5, 1.0]], [[0.5, 0.0]], *SEG, 0.01)
    assert added == 1 and A2.shape == (5, 2)
    np.testing.assert_array_equal(A2[4], [0.0, 1.0])
    assert b2[4] == np.float32(0.99) and float(b2[4]) >= 0.99
    np.testing.assert_array_equal(A2[:4], A)
    # oblique: a = (3, 4)/5, b = a.scol - margin = (0.6*3 + 0.8*4) - 0.01 = 4.99
    A8, b8 = _box([-8, -8], [8, 8])
    A2, b2, _ = orc.edit_region(A8, b8, [[3.0, 4.0]], [[0.0, 0.0]], [-1, 0], [1, 0], 0.01)
    np.testing.assert_allclose(A2[-1], [0.6, 0.8], rtol=0, atol=1e-7)
    assert abs(float(b2[-1]) - 4.99) < 1e-6 and float(b2[-1]) >= 4.99 - 1e-7


def test_face_relaxed_to_keep_segment(orc):
    A, b = _box([-2, -2], [2, 2])
    # the segment end (1, 0.995) lies beyond y = 0.99: the face moves up by
    # 0.005 + 2e-6 so that the whole segment stays inside
    A2, b2, added = orc.edit_region(A, b, [[0.0, 1.0]], [[0.0, 0.0]], [-1, 0], [1, 0.995], 0.01)
    assert added == 1
    want = 0.995 + 2e-6
    assert want <= float(b2[4]) < want + 2e-7


def test_order_and_redundancy(orc):
    A, b = _box([-2, -2], [2, 2])
    # the nearer collision (listed second) is placed first; its face y <= 0.49
    # leaves the farther one (0, 1) outside, so only one face is added
    A2, b2, added = orc.edit_region(A, b, [[0.0, 1.0], [0.0, 0.5]], [[0.0, 0.0], [0.0, 0.0]],
                                    *SEG, 0.01)
    assert added == 1 and b2[4] == np.float32(0.49)
    # a collision on the other side stays inside the first face: two faces,
    # nearest first
    A2, b2, added = orc.edit_region(A, b, [[0.0, -1.0], [0.0, 0.5]], [[0.0, 0.0], [0.0, 0.0]],
                                    *SEG, 0.01)
    assert added == 2
    np.testing.assert_array_equal(A2[4:], [[0.0, 1.0], [0.0, -1.0]])
    np.testing.assert_array_equal(b2[4:], np.float32([0.49, 0.99]))
    # equal distances: the lower index goes first
    A2, b2, added = orc.edit_region(A, b, [[1.0, 0.0], [0.0, 1.0]], [[0.0, 0.0], [0.0, 0.0]],
                                    [0, 0], [0, 0], 0.01)
    assert added == 2
    np.testing.assert_array_equal(A2[4:], [[1.0, 0.0], [0.0, 1.0]])


def test_points_outside_or_on_the_segment(orc):
    A, b = _box([-2, -2], [2, 2])
    # outside the region: no face
    A2, b2, added = orc.edit_region(A, b, [[3.0, 0.5]], [[1.0, 0.0]], *SEG, 0.01)
    assert added == 0 and A2.shape == A.shape
    # on the boundary (A.x - b == 0) is not strictly inside: no face
    _, _, added = orc.edit_region(A, b, [[2.0, 0.5]], [[1.0, 0.0]], *SEG, 0.01)
    assert added == 0
    # on the segment itself: a zero direction stays zero (Eigen's normalize),
    # and b is raised to 2e-6, a face every point satisfies
    A2, b2, added = orc.edit_region(A, b, [[0.5, 0.0]], [[0.5, 0.0]], *SEG, 0.01)
    assert added == 1
    np.testing.assert_array_equal(A2[4], [0.0, 0.0])
    assert 2e-6 <= float(b2[4]) < 2.01e-6


# The reference's test: three segments through the free cross between four
# corner boxes, four trajectory collisions repeated for each segment, three
# voxels (cuda_edit_regions_test.cpp:74-124).
VOX = np.array([[-0.55, -1.50, 1.90], [2.02, 1.80, 1.90], [0.00, 0.00, 0.00]], np.float32)
LS = np.array([[-1.90, 0.00, 0.00], [0.00, 0.00, 1.80]], np.float32)
LE = np.array([[0.00, 0.00, 1.90], [0.00, 1.80, 1.80]], np.float32)
COL = np.array([[-0.9, 0.6, 1.1, 1.9], [1.0, 0.5, -1.0, 1.9]], np.float32)


def reference_case():
    p = parse("boxes_in_corners.urdf")
    plant = p.build_plant()
    collisions = np.tile(COL, (1, 3))
    seg = [0] * 4 + [1] * 4 + [2] * 4
    return plant, collisions, seg


def test_reference_scenario(orc):
    plant, collisions, seg = reference_case()
    mp = plant.getMinimalPlant()
    op = orc.Plant(mp)
    rg = plant.getSceneInspector().robot_geometry_ids
    A, b = _box(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    # every collision is one: three inside corner boxes, (1.9, 1.9) on a voxel
    assert not orc.check_voxels(op, COL, VOX, 0.05, rg).any()
    regions, proj, opt = orc.edit_regions(op, collisions, seg, LS, LE, [(A, b)] * 3, rg, VOX,
                                          0.05, margin=0.01, steps=9)
    for r, (Ar, br) in enumerate(regions):
        assert _in_set(Ar, br, LS[:, r]) and _in_set(Ar, br, LE[:, r])
        for c in range(opt.shape[1]):
            assert not _in_set(Ar, br, opt[:, c])
        for c in range(COL.shape[1]):
            assert not _in_set(Ar, br, COL[:, c])
        assert Ar.shape[0] > 4
    # optimised collisions still collide and lie within 2^-9 of the bisection
    # interval's length from a free point on the way to the projection
    assert not orc.check_voxels(op, opt, VOX, 0.05, rg).any()
    # (1.9, 1.9) against segment 2 (y = 1.8): contact with the voxel sphere
    # (radius 0.05 + 0.01) at y = 1.84, approached from the free side
    y = opt[1, 11]
    assert opt[0, 11] == np.float32(1.9) and 1.84 <= y <= 1.84 + 0.1 / 512 + 1e-6
    # one bisection step whose midpoint is free leaves the collision unchanged
    _, _, opt1 = orc.edit_regions(op, collisions, seg, LS, LE, [(A, b)] * 3, rg, VOX, 0.05,
                                  margin=0.01, steps=1)
    assert np.linalg.norm(opt1[:, 1] - COL[:, 1]) <= 1e-6


def test_grouping_follows_segment_order(orc):
    plant, collisions, seg = reference_case()
    op = orc.Plant(plant.getMinimalPlant())
    rg = plant.getSceneInspector().robot_geometry_ids
    A, b = _box(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    perm = np.random.default_rng(3).permutation(12)
    r1, p1, o1 = orc.edit_regions(op, collisions, seg, LS, LE, [(A, b)] * 3, rg, VOX, 0.05)
    r2, p2, o2 = orc.edit_regions(op, collisions[:, perm], np.asarray(seg)[perm], LS, LE,
                                  [(A, b)] * 3, rg, VOX, 0.05)
    # grouped by segment, order of appearance inside a group
    for s in range(3):
        idx = perm[np.asarray(seg)[perm] == s]
        np.testing.assert_array_equal(o2[:, 4 * s:4 * s + 4],
                                      o1[:, idx])
    for (a1, b1), (a2, b2) in zip(r1, r2):
        assert a1.shape == a2.shape  # same faces, independent of input order
        np.testing.assert_array_equal(np.sort(b1), np.sort(b2))


@pytest.mark.parametrize("steps", [0, 3])
def test_no_step_keeps_collisions(orc, steps):
    plant, collisions, seg = reference_case()
    op = orc.Plant(plant.getMinimalPlant())
    rg = plant.getSceneInspector().robot_geometry_ids
    A, b = _box(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    _, proj, opt = orc.edit_regions(op, collisions, seg, LS, LE, [(A, b)] * 3, rg, VOX, 0.05,
                                    steps=steps)
    if steps == 0:
        np.testing.assert_array_equal(opt, collisions)
    # projections lie on their segments
    for c in range(12):
        s = seg[c]
        d = LE[:, s] - LS[:, s]
        t = np.dot(proj[:, c] - LS[:, s], d) / np.dot(d, d)
        np.testing.assert_allclose(proj[:, c], LS[:, s] + t * d, atol=1e-6)
        assert -1e-6 <= t <= 1 + 1e-6
