"""Pins the CPU restatement (oracle/) before it is trusted as the checker:
reference golden matrices and known-answer configurations, closed forms,
brute force, and published known-answer vectors."""
import math

import numpy as np
import pytest

from conftest import parse


def is_approx(a, b, prec=1e-5):
    """Eigen's isApprox: ||a - b|| <= prec * min(||a||, ||b||)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


# ---------------------------------------------------------------- FK golden
def test_fk_golden_2_branch_5dof(orc):
    # minimal_kinematic_tree_test.cpp:51-94, kinematic_tree_test.cpp:24-67
    p = parse("2_branch_5dof.urdf")
    pl = orc.Plant(p.get_minimal_plant())
    t = p.get_kinematic_tree()
    q = np.ones(5, np.float32)
    X = orc.fk(pl, q)
    Xd = orc.fk(pl, q.astype(np.float64), double=True)
    exp = {
        "world": np.eye(4),
        "link1": [[0.540302, -0.841471, 0, 0], [0.841471, 0.540302, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]],
        "link2": [[-0.41614684, -0.90929743, 0.0, 0.54030231], [0.90929743, -0.41614684, 0.0, 0.84147098],
                  [0, 0, 1, 0], [0, 0, 0, 1]],
        "link6": [[-0.9899925, -0.14112001, 0.0, -1.0596646], [0.14112001, -0.9899925, 0.0, 0.58067513],
                  [0, 0, 1, 0], [0, 0, 0, 1]],
    }
    for name, m in exp.items():
        i = t.get_link_index(name)
        assert is_approx(X[i], m), name
        assert is_approx(Xd[i], m), name


def test_fk_golden_prismatic(orc):
    # minimal_kinematic_tree_test.cpp:97-125, cuda_forward_kinematics_test.cpp:60-72
    p = parse("3dofrobot_prismatic.urdf")
    pl = orc.Plant(p.get_minimal_plant())
    i = p.get_kinematic_tree().get_link_index("link3")
    exp = [[0.60952558, 0.17919833, 0.77224771, 1.75935346],
           [0.59500984, 0.54030231, -0.59500984, 1.19001968],
           [-0.52387199, 0.8221687, 0.22270331, -1.88921496],
           [0, 0, 0, 1]]
    assert is_approx(orc.fk(pl, np.ones(3, np.float32))[i], exp)
    assert is_approx(orc.fk(pl, np.ones(3), double=True)[i], exp)


def test_fk_planar_chain_closed_form(orc):
    # a z-axis chain: link k+1 origin = sum_{i<=k} l_i (cos(sum q), sin(sum q))
    p = parse("2_branch_5dof.urdf")
    pl = orc.Plant(p.get_minimal_plant())
    t = p.get_kinematic_tree()
    rng = np.random.default_rng(0)
    for _ in range(20):
        q = rng.uniform(-3, 3, 5)
        X = orc.fk(pl, q, double=True)
        a1, a2 = q[0], q[0] + q[1]
        exp2 = [math.cos(a1), math.sin(a1)]
        exp3 = [math.cos(a1) + math.cos(a2), math.sin(a1) + math.sin(a2)]
        np.testing.assert_allclose(X[t.get_link_index("link2")][:2, 3], exp2, atol=1e-12)
        np.testing.assert_allclose(X[t.get_link_index("link3")][:2, 3], exp3, atol=1e-12)
        np.testing.assert_allclose(X[t.get_link_index("link3")][:2, :2],
                                   [[math.cos(q[0] + q[1] + q[2]), -math.sin(q[0] + q[1] + q[2])],
                                    [math.sin(q[0] + q[1] + q[2]), math.cos(q[0] + q[1] + q[2])]],
                                   atol=1e-12)


def test_fk_f32_matches_f64(orc):
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1000)
    op = orc.Plant(pl.getMinimalPlant())
    q = scenes.uniform_configs(7, 2000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    err = orc.fk_error(op, q)
    assert err.max() < 5e-6


# ---------------------------------------------------------------- pair rule
def test_pairs_2_branch_5dof():
    # urdf_parser_test.cpp:323-337
    p = parse("2_branch_5dof.urdf")
    expected = {(0, 3), (0, 5), (0, 6), (0, 7), (0, 8), (1, 3), (1, 5), (1, 6), (1, 7), (1, 8), (2, 4),
                (2, 5), (2, 6), (2, 7), (2, 8), (3, 4), (3, 5), (3, 6), (3, 7), (3, 8), (4, 6), (4, 7),
                (4, 8), (5, 7), (5, 8), (6, 7), (6, 8)}
    pairs = set(map(tuple, p.get_scene_inspector().robot_to_world_collision_pairs))
    assert expected <= pairs


def test_pair_counts_synthetic():
    from csdecomp_amd import scenes
    assert scenes.arm_scene(1000, 16, 0).getMinimalPlant().num_collision_pairs == 784
    assert scenes.arm_scene(1001, 48, 16).getMinimalPlant().num_collision_pairs == 2128
    assert scenes.dual_arm_scene(1002, 72, 24).getMinimalPlant().num_collision_pairs == 7056


# ---------------------------------------------------------------- known answers
KNOWN = [
    # collision_checker_test.cpp:28-105 (file, free configs, colliding configs)
    ("3dofrobot.urdf", [[0, 1.8, 2.46]], [[0, 1.8, 2.47]]),
    ("2_branch_5dof.urdf", [[0.6714, -0.9096, 0.79, -0.1806, 0.], [0.0204, 0.3714, 0.4714, 1.0344, -1.2816]],
     [[0.6714, -0.9096, 0.78, -0.1806, 0.], [0.0204, 0.3714, 0.4714, 1.0874, -1.2816]]),
    ("3dofrobot_spheres.urdf", [[0, 0, 2.67]], [[0, 0, 2.68]]),
    ("movable_block.urdf",
     np.array([[-0.375, -0.375, 0.004, 0.004, -0.003, -0.002],
               [-0.622, -0.416, 0.002, 0.04, 0.088, 0.303],
               [-0.099, -0.099, 0.0, 0.008, 0.008, 0.692],
               [1.0224, -0.3276, 0.0014, -0.0056, -0.8796, 0.0564],
               [0.7724, 0.7724, 0.0024, 0.0024, -0.7956, -0.7956]]).T,
     np.array([[-0.375, -0.375, -0.002, -0.002, 0.237, 0.591],
               [-0.461, -0.416, 0.088, 0.088, 0.202, 0.318],
               [-0.099, -0.099, 0.008, 0.008, 0.592, 0.449],
               [1.0224, -0.8546, -0.8796, 0.0564, -0.0256, -0.0256],
               [0.7724, 0.7974, -0.7956, -0.7956, -0.7956, -0.0116]]).T),
]


@pytest.mark.parametrize("fname,free,col", KNOWN, ids=[k[0] for k in KNOWN])
@pytest.mark.parametrize("rule", [0, 1], ids=["cpu_rule", "gpu_rule"])
def test_known_answer_configs(orc, fname, free, col, rule):
    pl = orc.Plant(parse(fname).get_minimal_plant())
    free = np.asarray(free, np.float32).T
    col = np.asarray(col, np.float32).T
    assert np.all(orc.check(pl, free, rule) == 1)
    assert np.all(orc.check(pl, col, rule) == 0)
    assert np.all(orc.min_sd(pl, free) > 0)
    assert np.all(orc.min_sd(pl, col) < 0)


def voxel_grid_6x6x6():
    # collision_checker_test.cpp:107-140
    nx, ny, nz, spx, spy, spz = 6, 6, 6, 0.3, 0.3, 0.1
    off = np.array([-(nx - 1) * spx / 2, -(ny - 1) * spy / 2, 0], np.float32)
    pts = [np.array([ix * spx, iy * spy, iz * spz], np.float32) + off
           for ix in range(nx) for iy in range(ny) for iz in range(nz)]
    return np.asarray(pts, np.float32).T


VOX_FREE = np.array([[0.00, 0.00, 0.00], [0.00, 0.00, 0.00], [0.00, 0.00, 0.27], [0.00, 0.00, 0.57],
                     [0.00, 1.50, 1.27]], np.float32)
VOX_COL = np.array([[0.15, 0.00, 0.00, 0.00, 0.16], [0.00, 0.15, 0.48, 0.48, 0.28],
                    [0.00, 0.00, 0.00, 0.31, 0.06], [0.00, 0.87, -0.43, -0.43, 1.46],
                    [0.00, 1.17, 1.17, 1.17, 1.59]], np.float32)


@pytest.mark.parametrize("rule", [0, 1], ids=["cpu_rule", "gpu_rule"])
def test_known_answer_voxels(orc, rule):
    p = parse("movable_block.urdf")
    pl = orc.Plant(p.get_minimal_plant())
    rg = list(p.get_scene_inspector().robot_geometry_ids)
    vox = voxel_grid_6x6x6()
    assert np.all(orc.check_voxels(pl, VOX_FREE, vox, 0.02, rg, rule) == 1)
    assert np.all(orc.check_voxels(pl, VOX_COL, vox, 0.02, rg, rule) == 0)


def test_sphere_sphere_closed_form(orc):
    # 3dofrobot_spheres: the q3 sweep crosses contact where the closed form
    # distance between the two sphere centres equals r1 + r2
    pl = orc.Plant(parse("3dofrobot_spheres.urdf").get_minimal_plant())
    qs = np.linspace(2.60, 2.75, 301).astype(np.float32)
    q = np.stack([np.zeros_like(qs), np.zeros_like(qs), qs])
    sd = orc.min_sd(pl, q)
    f = orc.check(pl, q, 1)
    outside = np.abs(sd) >= 1e-5
    assert np.all((sd > 0)[outside] == (f[outside] == 1))
    crossing = qs[np.argmax(sd < 0)]
    assert 2.67 < crossing <= 2.68


def test_signed_distance_sign_matches_flags(orc):
    # the sign of the FP64 signed distance agrees with both comparison rules'
    # flags away from contact on the benchmark scene (its magnitude is pinned
    # by brute force in test_oracle_sd.py)
    from csdecomp_amd import scenes
    rng = np.random.default_rng(3)
    pl = scenes.arm_scene(1000)
    op = orc.Plant(pl.getMinimalPlant())
    q = scenes.uniform_configs(11, 3000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    sd = orc.min_sd(op, q)
    for rule in (0, 1):
        f = orc.check(op, q, rule)
        out = np.abs(sd) >= 1e-5
        assert np.all((sd > 0)[out] == (f[out] == 1))
    assert (sd > 0).sum() > 10 and (sd < 0).sum() > 10


def test_flags_rules_agree_off_band(orc):
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    op = orc.Plant(pl.getMinimalPlant())
    q = scenes.uniform_configs(2001, 4000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    sd = orc.min_sd(op, q)
    a, b = orc.check(op, q, 0), orc.check(op, q, 1)
    out = np.abs(sd) >= 1e-5
    assert np.array_equal(a[out], b[out])


# ---------------------------------------------------------------- edges
def test_edge_rule(orc):
    # checkEdgeCollisionFree: ceil(d/step) points, q1 included, q2 excluded
    pl = orc.Plant(parse("3dofrobot_spheres.urdf").get_minimal_plant())
    q1 = np.array([0, 0, 2.0], np.float32)
    q2 = np.array([0, 0, 2.675], np.float32)  # endpoint free, crossing beyond
    assert orc.edge_num_steps(q1, q2, 0.01) == math.ceil(np.float32(0.675) / np.float32(0.01))
    assert orc.check_edge(pl, q1, q2, 0.01)
    q3 = np.array([0, 0, 2.9], np.float32)
    assert not orc.check_edge(pl, q1, q3, 0.01)
    assert orc.check_edge(pl, q1, q1, 0.01)  # zero-length edge checks nothing


# ---------------------------------------------------------------- RNG
def test_philox_known_answers(orc):
    # Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11)
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, exp in kat:
        assert list(orc.philox(ctr, key)) == exp


# ---------------------------------------------------------------- sampler
def test_hit_and_run_invariants(orc):
    # cuda_hit_and_run_sampling_test.cpp:106-154: samples stay inside and move
    A = np.array([[-1, 0], [1, 0], [0, -1], [0, 1], [1, 1]], np.float32)
    b = np.array([0, 6, 0, 6, 11.25], np.float32)
    s = orc.hit_and_run(A, b, [1, 1], 20000, 50, seed=1337)
    assert np.all(s @ A.T <= b + 1e-4)
    assert np.unique(s[:, 0]).size > 19000
    # coverage of the acute corner (5.75, 5.5) (cuda_hit_and_run_sampling_test.cpp:183-197)
    assert np.min(np.linalg.norm(s - [5.75, 5.5], axis=1)) < 0.05


def test_hit_and_run_uniform_on_box(orc):
    A = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], np.float32)
    b = np.array([1, 1, 2, 2, 0.5, 0.5], np.float32)
    s = orc.hit_and_run(A, b, [0, 0, 0], 40000, 40, seed=7)
    np.testing.assert_allclose(s.mean(0), 0, atol=0.03)
    np.testing.assert_allclose(s.var(0), np.array([1, 4, 0.25]) / 3.0, rtol=0.05)
    # chi-square over a 4x4x4 grid
    h, _ = np.histogramdd(s, bins=4, range=[(-1, 1), (-2, 2), (-0.5, 0.5)])
    exp = s.shape[0] / 64
    chi2 = ((h - exp) ** 2 / exp).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5


# ---------------------------------------------------------------- EI-ZO stages
def test_num_samples_unadaptive(orc):
    # M_0 = 27,948 at eps 1e-3, delta 0.05, tau 0.5 (closed form)
    d0 = orc.delta_k(0.05, 0)
    assert d0 == np.float32(np.float32(6 * np.float32(0.05)) / (math.pi ** 2))
    assert orc.num_samples_unadaptive(1e-3, d0, 0.5) == 27948
    assert math.ceil(2 * math.log(1 / d0) / (1e-3 * 0.25)) == 27948
    # later rounds need more samples (delta_k shrinks as 1/k^2)
    Ms = [orc.num_samples_unadaptive(1e-3, orc.delta_k(0.05, k), 0.5) for k in range(10)]
    assert all(a < b for a, b in zip(Ms, Ms[1:]))


def test_projection_reference_values(orc):
    # cuda_set_builder_utils_test.cpp:18-68
    q1, q2 = np.array([0, 1], np.float32), np.array([2, 1.8], np.float32)
    p0, d0 = orc.project(np.array([0, 0], np.float32), q1, q2)
    np.testing.assert_allclose(p0, q1, atol=1e-6)
    assert abs(d0 - 1.0) < 1e-6
    p1, d1 = orc.project(np.array([3.5, 0], np.float32), q1, q2)
    np.testing.assert_allclose(p1, q2, atol=1e-6)
    assert abs(d1 - np.linalg.norm([3.5 - 2, -1.8])) < 1e-6
    p2, d2 = orc.project(np.array([1.5, 0], np.float32), q1, q2)
    assert np.linalg.norm(p2 - [0.948276, 1.37931]) <= 1e-6
    assert abs(np.linalg.norm(p2 - [1.5, 0]) - d2) < 1e-5


def test_place_faces_closed_form(orc):
    # one particle: face normal = (scol - proj)/|.|, offset = a.scol - margin,
    # relaxed to keep the segment inside
    cap, dim = 8, 2
    A = np.zeros((cap, dim), np.float32)
    b = np.zeros(cap, np.float32)
    scol = np.array([[0.5, 1.0], [0.5, 1.5], [0.2, -0.9]], np.float32)
    proj = np.array([[0.5, 0.0], [0.5, 0.0], [0.2, 0.0]], np.float32)
    ls, le = np.array([0, 0], np.float32), np.array([1, 0], np.float32)
    added, nf = orc.place_faces(A, b, 0, cap, scol, proj, ls, le, 0.01, 10)
    # particle 0 is closest (d=1.0 vs 1.5 vs 0.9 -> particle 2 first)
    assert added == 2 and nf == 2
    np.testing.assert_allclose(A[0], [0, -1], atol=1e-7)
    assert abs(b[0] - (0.9 - 0.01)) < 1e-6
    np.testing.assert_allclose(A[1], [0, 1], atol=1e-7)
    assert abs(b[1] - (1.0 - 0.01)) < 1e-6  # particle 1 made redundant by face 2
    # relaxation: a face through the segment is pushed out to contain it
    A2 = np.zeros((cap, dim), np.float32)
    b2 = np.zeros(cap, np.float32)
    added, nf = orc.place_faces(A2, b2, 0, cap, np.array([[2.0, 0.0]], np.float32),
                                np.array([[1.0, 0.0]], np.float32), ls, le, 0.01, 10)
    assert added == 1 and b2[0] >= 1.0 and A2[0] @ le <= b2[0]


def test_oracle_eizo_boxes_in_corners(orc):
    # FCItest1 (edge_inflation_zero_order_test.cpp:118-172) with the CPU
    # restatement: endpoints inside, (1.8, 0.2) inside, (1.8, 0.4) outside
    from csdecomp_amd import scenes
    p = scenes.boxes_in_corners()
    mp = p.get_minimal_plant()
    pl = orc.Plant(mp)
    t = p.get_kinematic_tree()
    lo, hi = t.get_position_lower_limits(), t.get_position_upper_limits()
    A = np.zeros((4, 2), np.float32)
    bdom = np.zeros(4, np.float32)
    for i in range(2):
        A[2 * i, i], A[2 * i + 1, i] = -1, 1
        bdom[2 * i], bdom[2 * i + 1] = -lo[i], hi[i]
    opts = (0.01, 0.5, 0.05, 0.01, 9, 1000, 10, 500, 50, 1337)
    ls, le = np.array([0, 0], np.float32), np.array([1, 0], np.float32)
    rg = list(p.get_scene_inspector().robot_geometry_ids)
    Ar, br, info = orc.eizo(pl, rg, A, bdom, ls, le, opts)
    assert info.converged
    inside = lambda x: np.all(Ar @ np.asarray(x, np.float32) <= br + 1e-6)
    assert inside(ls) and inside(le)
    assert inside([1.8, 0.2]) and not inside([1.8, 0.4])
    # certified: collision fraction of uniform samples below 1.1 eps
    s = orc.hit_and_run(Ar, br, (ls + le) / 2, 4000, 50, seed=99, stream=5)
    frac = 1 - orc.check(pl, s.T, 1).mean()
    assert frac < 0.011
def test_oracle_eizo_batch_equals_serial(orc):
    """orc_eizo_batch (regions side by side, one per OpenMP thread, the CPU
    baseline's path) returns, per region, exactly what orc_eizo returns for
    that region with stream r, in both the checking and the speed builds'
    wrappers (the speed build only where its arithmetic agrees: faces count
    and certification)."""
    from csdecomp_amd import scenes
    p = scenes.boxes_in_corners()
    mp = p.get_minimal_plant()
    pl = orc.Plant(mp)
    t = p.get_kinematic_tree()
    lo, hi = t.get_position_lower_limits(), t.get_position_upper_limits()
    A = np.zeros((4, 2), np.float32)
    bdom = np.zeros(4, np.float32)
    for i in range(2):
        A[2 * i, i], A[2 * i + 1, i] = -1, 1
        bdom[2 * i], bdom[2 * i + 1] = -lo[i], hi[i]
    opts = (0.01, 0.5, 0.05, 0.01, 9, 500, 10, 500, 30, 1337)
    ls = np.array([[0, 0], [0.2, -0.3], [-0.5, 0.1], [0.3, 0.3], [1.0, -0.2]], np.float32).T
    le = np.array([[1, 0], [0.2, 0.3], [-0.2, 0.1], [0.4, 0.2], [1.2, -0.1]], np.float32).T
    rg = list(p.get_scene_inspector().robot_geometry_ids)
    Ab, bb, nf, st, infos = orc.eizo_batch(pl, rg, A, bdom, ls, le, opts, cap=64, threads=3)
    for r in range(ls.shape[1]):
        Ar, br, info = orc.eizo(pl, rg, A, bdom, ls[:, r], le[:, r], opts, cap=64, stream=r)
        assert st[r] == 0 and nf[r] == len(br)
        np.testing.assert_array_equal(Ab[r, :nf[r]], Ar)
        np.testing.assert_array_equal(bb[r, :nf[r]], br)
        assert (infos[r].iterations, infos[r].converged, infos[r].samples_drawn) == \
            (info.iterations, info.converged, info.samples_drawn)
    _, _, nf2, st2, infos2 = orc.eizo_batch(pl, rg, A, bdom, ls, le, opts, cap=64, fast=True)
    assert (st2 == 0).all() and all(i.converged for i in infos2)


