"""GPU parity of FK and collision checking against the CPU restatement.

Bar: flags identical except configurations whose FP64 signed distance lies
within 1e-5 m of contact; FK within 1e-5 (FP32 vs FP64) per link."""
import numpy as np
import pytest

from conftest import assert_flags_match, parse
from test_oracle import KNOWN, VOX_COL, VOX_FREE, voxel_grid_6x6x6

pytestmark = pytest.mark.gpu


def _synthetic(name):
    from csdecomp_amd import scenes
    return {"arm7_16box": lambda: scenes.arm_scene(1000, 16, 0),
            "arm7_48box16sph": lambda: scenes.arm_scene(1001, 48, 16),
            "dualarm14_72box24sph": lambda: scenes.dual_arm_scene(1002, 72, 24)}[name]()


@pytest.mark.parametrize("fname,free,col", KNOWN, ids=[k[0] for k in KNOWN])
def test_known_answer_configs_api(gpu, csd, fname, free, col):
    mp = parse(fname).get_minimal_plant()
    free = np.asarray(free, np.float32).T
    col = np.asarray(col, np.float32).T
    assert list(csd.CheckCollisionFreeCuda(free, mp)) == [1] * free.shape[1]
    assert list(csd.CheckCollisionFreeCuda(col, mp)) == [0] * col.shape[1]
    for k in range(free.shape[1]):
        assert csd.CheckCollisionFree(free[:, k], mp)
    for k in range(col.shape[1]):
        assert not csd.CheckCollisionFree(col[:, k], mp)


def test_known_answer_voxels_api(gpu, csd):
    p = parse("movable_block.urdf")
    mp = p.get_minimal_plant()
    rg = list(p.get_scene_inspector().robot_geometry_ids)
    vox = csd.Voxels(voxel_grid_6x6x6())
    assert list(csd.CheckCollisionFreeVoxelsCuda(VOX_FREE, vox, 0.02, mp, rg)) == [1, 1, 1]
    assert list(csd.CheckCollisionFreeVoxelsCuda(VOX_COL, vox, 0.02, mp, rg)) == [0] * 5


PLANTS = ["2_branch_5dof.urdf", "3dofrobot.urdf", "3dofrobot_prismatic.urdf",
          "3dofrobot_spheres.urdf", "movable_block.urdf", "arm7_16box", "arm7_48box16sph", "dualarm14_72box24sph"]


def _plant(name):
    if name.endswith(".urdf"):
        p = parse(name)
        t = p.get_kinematic_tree()
        lo = np.asarray(t.get_position_lower_limits())
        hi = np.asarray(t.get_position_upper_limits())
        return p.get_minimal_plant(), np.clip(lo, -4, 4), np.clip(hi, -4, 4)
    pl = _synthetic(name)
    return pl.getMinimalPlant(), pl.getPositionLowerLimits(), pl.getPositionUpperLimits()


@pytest.mark.parametrize("name", PLANTS)
def test_random_flags_vs_oracle(gpu, csd, orc, name):
    import torch
    from csdecomp_amd import capi, scenes
    mp, lo, hi = _plant(name)
    n = 20000 if name != "dualarm14_72box24sph" else 8000
    q = scenes.uniform_configs(2000 + len(name), n, lo, hi)  # (dof, n)
    op = orc.Plant(mp)
    sd = orc.min_sd(op, q)
    ref = orc.check(op, q, orc.RULE_GPU)
    # reference API path: (dim, N) column-major host array
    api = np.asarray(csd.CheckCollisionFreeCuda(q, mp), np.uint8)
    assert_flags_match(api, ref, sd, f"{name} api")
    # C-ABI path, structure-of-arrays on the device
    dp = capi.DevicePlant(mp)
    qs = torch.from_numpy(np.ascontiguousarray(q)).to(gpu)  # [dof][n]
    out = torch.empty(n, dtype=torch.uint8, device=gpu)
    capi.check_collision_free(dp, qs, n, 1, n, out)
    torch.cuda.synchronize()
    assert_flags_match(out.cpu().numpy(), ref, sd, f"{name} soa")


def test_fast_path_selected(gpu):
    from csdecomp_amd import capi
    for name in ("arm7_16box", "arm7_48box16sph", "dualarm14_72box24sph"):
        mp, _, _ = _plant(name)
        assert capi.DevicePlant(mp).info()["fast_path"], name
    mp, _, _ = _plant("2_branch_5dof.urdf")
    assert not capi.DevicePlant(mp).info()["fast_path"]


def test_fk_golden_gpu(gpu, csd):
    from test_oracle import is_approx
    p = parse("3dofrobot_prismatic.urdf")
    tree = p.get_minimal_kinematic_tree()
    tf = csd.computeForwardKinematicsCuda(np.ones((3, 1), np.float32), tree)
    i = p.get_kinematic_tree().get_link_index("link3")
    exp = [[0.60952558, 0.17919833, 0.77224771, 1.75935346],
           [0.59500984, 0.54030231, -0.59500984, 1.19001968],
           [-0.52387199, 0.8221687, 0.22270331, -1.88921496], [0, 0, 0, 1]]
    assert is_approx(tf[:, 4 * i:4 * i + 4], exp)
    m = csd.computeLinkFrameToWorldTransformsMinimal(np.ones(3, np.float32), tree)
    assert is_approx(m[4 * i:4 * i + 4], exp)
    p2 = parse("2_branch_5dof.urdf")
    t2 = p2.get_kinematic_tree()
    tfs = t2.compute_link_frame_to_world_transforms(np.ones(5, np.float32))
    assert is_approx(tfs[t2.get_link_index("link6")],
                     [[-0.9899925, -0.14112001, 0.0, -1.0596646], [0.14112001, -0.9899925, 0.0, 0.58067513],
                      [0, 0, 1, 0], [0, 0, 0, 1]])


@pytest.mark.parametrize("name", ["2_branch_5dof.urdf", "3dofrobot_prismatic.urdf", "arm7_16box", "dualarm14_72box24sph"])
def test_fk_error_per_link(gpu, csd, orc, name):
    """FP32 GPU FK against the FP64 oracle, per link (max abs over the 3x4)."""
    from csdecomp_amd import scenes
    mp, lo, hi = _plant(name)
    q = scenes.uniform_configs(5, 500, lo, hi)
    tf = np.asarray(csd.computeForwardKinematicsCuda(q, mp.kin_tree))  # (4, 4*L*N)
    op = orc.Plant(mp)
    L = mp.kin_tree.num_links
    worst = np.zeros(L)
    for k in range(q.shape[1]):
        Xd = orc.fk(op, q[:, k].astype(np.float64), double=True)
        for l in range(L):
            g = tf[:, 4 * (k * L + l):4 * (k * L + l) + 4]
            worst[l] = max(worst[l], np.abs(g[:3, :] - Xd[l][:3, :]).max())
    assert worst.max() < 1e-5, worst


def test_voxels_vs_oracle(gpu, csd, orc):
    from csdecomp_amd import scenes
    p = parse("movable_block.urdf")
    mp = p.get_minimal_plant()
    rg = list(p.get_scene_inspector().robot_geometry_ids)
    t = p.get_kinematic_tree()
    q = scenes.uniform_configs(9, 3000, t.get_position_lower_limits(), t.get_position_upper_limits())
    vox = voxel_grid_6x6x6()
    op = orc.Plant(mp)
    ref = orc.check_voxels(op, q, vox, 0.02, rg, orc.RULE_GPU)
    sd = orc.min_sd(op, q, vox, 0.02, rg)
    got = np.asarray(csd.CheckCollisionFreeVoxelsCuda(q, csd.Voxels(vox), 0.02, mp, rg), np.uint8)
    assert_flags_match(got, ref, sd, "voxels")
    flags = csd.CheckCollisionFreeVoxelsWithoutSelfCollisionCuda(q[:, :200], csd.Voxels(vox), 0.02, mp, rg)
    ref2 = orc.voxel_flags(op, q[:, :200], vox, 0.02, rg, orc.RULE_GPU)
    # a (configuration, voxel) flag may differ only where the robot is within
    # 1e-5 m of that voxel's sphere: there a voxel radius 1e-5 m smaller and
    # one 1e-5 m larger give the oracle different verdicts
    lo = orc.voxel_flags(op, q[:, :200], vox, 0.02 - 1e-5, rg, orc.RULE_GPU)
    hi = orc.voxel_flags(op, q[:, :200], vox, 0.02 + 1e-5, rg, orc.RULE_GPU)
    mism = np.asarray(flags) != ref2
    assert not np.any(mism & (lo == hi)), np.argwhere(mism & (lo == hi))[:5]
    assert mism.sum() <= 2, np.argwhere(mism)[:5]


def test_fast_path_voxels(gpu, csd, orc):
    """Robot spheres x voxels through the fast kernel."""
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1000, 16, 0)
    mp = pl.getMinimalPlant()
    rg = list(pl.getRobotGeometryIds())
    rng = np.random.default_rng(4)
    vox = rng.uniform(-1, 1, size=(3, 50)).astype(np.float32)
    q = scenes.uniform_configs(12, 5000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    op = orc.Plant(mp)
    ref = orc.check_voxels(op, q, vox, 0.03, rg, orc.RULE_GPU)
    sd = orc.min_sd(op, q, vox, 0.03, rg)
    got = np.asarray(csd.CheckCollisionFreeVoxelsCuda(q, csd.Voxels(vox), 0.03, mp, rg), np.uint8)
    assert_flags_match(got, ref, sd, "fast voxels")


def _capsule_scene():
    """The 7-DOF arm with 8 boxes and 6 capsule obstacles paired with every
    robot sphere: (scene, plant, geometries, pairs, index of the first capsule)."""
    from csdecomp_amd import scenes
    from csdecomp_amd import _pycsdecomp as c
    pl = scenes.arm_scene(1000, 8, 0)
    mp = pl.getMinimalPlant()
    geoms = list(mp.scene_collision_geometries)
    obs_link = geoms[-1].link_index
    rng = np.random.default_rng(5)
    start = len(geoms)
    for _ in range(6):
        g = c.CollisionGeometry()
        g.type = c.ShapePrimitive.CAPSULE
        r, ln = rng.uniform(0.05, 0.15), rng.uniform(0.2, 0.6)
        g.dimensions = np.array([r, ln, r], np.float32)
        g.link_index = obs_link
        g.group = c.CollisionGroup.WORLD
        X = np.eye(4, dtype=np.float32)
        X[:3, :3] = scenes.random_rotation(rng)
        X[:3, 3] = rng.uniform(-0.8, 0.8, 3)
        g.X_L_B = X
        geoms.append(g)
    robot = list(pl.getRobotGeometryIds())
    pairs = list(np.asarray(mp.collision_pairs_flat).reshape(-1, 2))
    for gi in range(start, len(geoms)):
        pairs += [(r, gi) for r in robot]
    mp.set_scene_collision_geometries(geoms)
    mp.set_collision_pairs_from_matrix(np.asarray(pairs, np.int32).T)
    return pl, mp, geoms, pairs, start


def test_capsules_vs_oracle(gpu, csd, orc):
    """Capsule obstacles (fast path) and a capsule link (generic path) against
    the FP64 segment distances of the oracle."""
    from csdecomp_amd import scenes
    from csdecomp_amd import _pycsdecomp as c
    pl, mp, geoms, pairs, start = _capsule_scene()
    q = scenes.uniform_configs(13, 6000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    op = orc.Plant(mp)
    sd = orc.min_sd(op, q)
    got = np.asarray(csd.CheckCollisionFreeCuda(q, mp), np.uint8)
    assert_flags_match(got, orc.check(op, q, orc.RULE_GPU), sd, "capsule obstacles")
    # a capsule riding on link 3 -> generic kernel (capsule-box, capsule-sphere)
    g = c.CollisionGeometry()
    g.type = c.ShapePrimitive.CAPSULE
    g.dimensions = np.array([0.04, 0.3, 0.04], np.float32)
    g.link_index = geoms[8].link_index
    g.group = c.CollisionGroup.ROBOT
    g.X_L_B = np.eye(4, dtype=np.float32)
    geoms.append(g)
    ci = len(geoms) - 1
    pairs += [(ci, gi) for gi in range(start - 8, len(geoms) - 1) if geoms[gi].group == c.CollisionGroup.WORLD]
    mp.set_scene_collision_geometries(geoms)
    mp.set_collision_pairs_from_matrix(np.asarray(pairs, np.int32).T)
    op = orc.Plant(mp)
    sd = orc.min_sd(op, q)
    got = np.asarray(csd.CheckCollisionFreeCuda(q, mp), np.uint8)
    assert_flags_match(got, orc.check(op, q, orc.RULE_GPU), sd, "capsule link")


@pytest.mark.parametrize("name", ["arm7_16box", "arm7_48box16sph", "dualarm14_72box24sph",
                                  "capsules"])
def test_bound_cull_whole_waves(gpu, csd, orc, name):
    """Each wave of 64 lanes holds copies of one configuration, so the
    bounding-sphere cull of the obstacle tests decides for the whole wave: a
    body or obstacle bound that missed a sphere, a box corner or a capsule
    end would drop that configuration's collisions. Configurations free of
    self collision are taken, so that the obstacle tests decide them."""
    import torch
    from csdecomp_amd import capi, scenes
    from csdecomp_amd import _pycsdecomp as c

    def plant():
        if name == "capsules":
            pl, mp, *_ = _capsule_scene()
            return mp, pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
        return _plant(name)

    mp, lo, hi = plant()
    # the same plant with only its robot x robot pairs
    ms, _, _ = plant()
    robot = {i for i, g in enumerate(ms.scene_collision_geometries)
             if g.group == c.CollisionGroup.ROBOT}
    pairs = np.asarray(ms.collision_pairs_flat).reshape(-1, 2)
    ms.set_collision_pairs_from_matrix(
        np.ascontiguousarray(pairs[[a in robot and b in robot for a, b in pairs]].T, np.int32))
    op = orc.Plant(mp)
    q = scenes.uniform_configs(77, 8000, lo, hi)
    self_free = orc.check(orc.Plant(ms), q, orc.RULE_GPU).astype(bool)
    q = np.ascontiguousarray(q[:, self_free][:, :1024])
    ref = orc.check(op, q, orc.RULE_GPU)
    sd = orc.min_sd(op, q)
    m = q.shape[1]
    assert m >= 256 and 0.05 < np.mean(ref) < 0.95, "want both outcomes among self-free configurations"
    qq = torch.from_numpy(np.ascontiguousarray(np.repeat(q, 64, axis=1))).to(gpu)
    out = torch.empty(m * 64, dtype=torch.uint8, device=gpu)
    dp = capi.DevicePlant(mp)
    for shape in (1, 2, 4, 11, 14):
        out.fill_(7)
        assert capi.tune_check(dp, qq, m * 64, 1, m * 64, out, force=shape) == shape
        capi.check_collision_free(dp, qq, m * 64, 1, m * 64, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(m, 64)
        assert np.all(got == got[:, :1]), f"shape {shape}: copies of one configuration disagree"
        assert_flags_match(got[:, 0], ref, sd, f"{name} whole waves, shape {shape}")


def test_pair_collision_free(gpu, csd):
    c = csd
    a, b = c.CollisionGeometry(), c.CollisionGeometry()
    a.type = b.type = c.ShapePrimitive.SPHERE
    a.dimensions = np.array([0.5, -1, -1], np.float32)
    b.dimensions = np.array([0.25, -1, -1], np.float32)
    Xa, Xb = np.eye(4, dtype=np.float32), np.eye(4, dtype=np.float32)
    Xb[0, 3] = 0.76
    assert c.PairCollisionFree(a, Xa, b, Xb)
    Xb[0, 3] = 0.74
    assert not c.PairCollisionFree(a, Xa, b, Xb)


def test_empty_and_errors(gpu, csd):
    import torch
    from csdecomp_amd import capi
    mp, _, _ = _plant("arm7_16box")
    assert list(csd.CheckCollisionFreeCuda(np.zeros((7, 0), np.float32), mp)) == []
    with pytest.raises(ValueError):
        csd.CheckCollisionFreeCuda(np.zeros((6, 3), np.float32), mp)
    dp = capi.DevicePlant(mp)
    out = torch.empty(4, dtype=torch.uint8, device=gpu)
    with pytest.raises(capi.CsdError) as e:
        capi.check_collision_free(dp, None, 4, 1, 4, out)
    assert e.value.status == 1
    with pytest.raises(capi.CsdError):
        capi.check_collision_free(dp, out, -1, 1, 4, out)


def test_split_checker_and_tuner(gpu, orc):
    """Batches that share each 64-configuration group between 2 or 4 waves
    (chosen by size or by the autotuner) give the oracle's flags."""
    import torch
    from csdecomp_amd import capi, scenes
    from conftest import assert_flags_match
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    dp = capi.DevicePlant(mp)
    op = orc.Plant(mp)
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    for n in (140000, 30000):  # 2 waves per group, then 4, by the size rule
        q = scenes.uniform_configs(n, n, lo, hi)
        dq = torch.from_numpy(np.ascontiguousarray(q)).to(gpu)
        out = torch.empty(n, dtype=torch.uint8, device=gpu)
        capi.check_collision_free(dp, dq, n, 1, n, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        ref = orc.check(op, q, orc.RULE_GPU)
        assert_flags_match(got, ref, orc.min_sd(op, q), f"n={n}")
        s = capi.tune_check(dp, dq, n, 1, n, out)
        assert s in (1, 2, 4, 11, 12, 14, 20)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), got)
        capi.check_collision_free(dp, dq, n, 1, n, out)  # uses the tuned split
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), got)


@pytest.mark.parametrize("scene", ["arm16", "arm64", "dual"])
def test_two_pass_checker(gpu, orc, scene):
    """Every checker shape -- one pass with 1, 2, 4 waves per 64
    configurations and two passes (self pairs, then obstacles for the
    survivors) with 1, 2, 4 -- gives the same flags, equal to the oracle's
    outside the contact band; also with voxels and a ragged batch."""
    import torch
    from csdecomp_amd import capi, scenes
    from conftest import assert_flags_match
    pl = {"arm16": lambda: scenes.arm_scene(1000, 16, 0),
          "arm64": lambda: scenes.arm_scene(1001, 48, 16),
          "dual": lambda: scenes.dual_arm_scene(1002, 72, 24)}[scene]()
    mp = pl.getMinimalPlant()
    dp = capi.DevicePlant(mp)
    op = orc.Plant(mp)
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    n = 20011
    q = scenes.uniform_configs(5, n, lo, hi)
    dq = torch.from_numpy(np.ascontiguousarray(q)).to(gpu)
    out = torch.empty(n, dtype=torch.uint8, device=gpu)
    ref = orc.check(op, q, orc.RULE_GPU)
    sd = orc.min_sd(op, q)
    first = None
    for shape in (1, 2, 4, 11, 12, 14):
        out.fill_(7)
        assert capi.tune_check(dp, dq, n, 1, n, out, force=shape) == shape
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert set(np.unique(got)) <= {0, 1}
        assert_flags_match(got, ref, sd, f"{scene} shape {shape}")
        out.fill_(7)
        capi.check_collision_free(dp, dq, n, 1, n, out)  # the recorded shape
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), got)
        if first is None:
            first = got
        np.testing.assert_array_equal(got, first)
    # voxels ride in pass 2
    rng = np.random.default_rng(9)
    vox = torch.from_numpy(rng.uniform(-0.8, 0.8, (40, 3)).astype(np.float32)).to(gpu)
    rg = list(pl.getRobotGeometryIds())
    vref = orc.check_voxels(op, q, vox.cpu().numpy().T, 0.03, rg)
    vsd = orc.min_sd(op, q, vox.cpu().numpy().T, 0.03, rg)
    for shape in (1, 14):
        capi.tune_check(dp, dq, n, 1, n, out, force=shape)
        out.fill_(7)
        capi.check_collision_free_voxels(dp, dq, n, 1, n, vox, 40, 0.03, rg, out)
        torch.cuda.synchronize()
        assert_flags_match(out.cpu().numpy(), vref, vsd, f"{scene} voxels shape {shape}")
    with pytest.raises(capi.CsdError):
        capi.tune_check(dp, dq, n, 1, n, out, force=3)


@pytest.mark.parametrize("scene", ["arm16", "arm64", "dual"])
def test_screened_checker(gpu, orc, scene):
    """Screened checking (tuner shape 20): FK and a greedy subset of the pair
    tests for every configuration, then the full check of the undecided
    ones. Flags equal the one-pass checker's bit for bit (and the oracle's
    outside the contact band) for ragged, tiny and strided batches; voxel
    checks keep the unscreened path."""
    import torch
    from csdecomp_amd import capi, scenes
    from conftest import assert_flags_match
    pl = {"arm16": lambda: scenes.arm_scene(1000, 16, 0),
          "arm64": lambda: scenes.arm_scene(1001, 48, 16),
          "dual": lambda: scenes.dual_arm_scene(1002, 72, 24)}[scene]()
    mp = pl.getMinimalPlant()
    dp = capi.DevicePlant(mp)
    op = orc.Plant(mp)
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    n = 20011
    q = scenes.uniform_configs(6, n, lo, hi)
    dq = torch.from_numpy(np.ascontiguousarray(q)).to(gpu)
    out = torch.empty(n, dtype=torch.uint8, device=gpu)
    assert capi.tune_check(dp, dq, n, 1, n, out, force=1) == 1
    torch.cuda.synchronize()
    base = out.cpu().numpy()
    out.fill_(7)
    shape = capi.tune_check(dp, dq, n, 1, n, out, force=20)
    assert shape in (1, 20)
    if scene == "arm16":
        assert shape == 20  # 8 bodies x 4 sphere slots: screened
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert set(np.unique(got)) <= {0, 1}
    np.testing.assert_array_equal(got, base)
    assert_flags_match(got, orc.check(op, q, orc.RULE_GPU), orc.min_sd(op, q), f"{scene} screened")
    out.fill_(7)
    capi.check_collision_free(dp, dq, n, 1, n, out)  # the recorded shape
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), base)
    # another batch of the same size class, and configurations as rows
    # (stride_n = dof, stride_j = 1)
    m = 16411
    qa = torch.from_numpy(np.ascontiguousarray(q[:, :m].T)).to(gpu)
    o2 = torch.full((m,), 7, dtype=torch.uint8, device=gpu)
    capi.check_collision_free(dp, qa, m, q.shape[0], 1, o2)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(o2.cpu().numpy(), base[:m])
    # batches smaller than one survivor workgroup's range
    for k in (1, 63, 700):
        ok = torch.full((k,), 7, dtype=torch.uint8, device=gpu)
        dk = torch.from_numpy(np.ascontiguousarray(q[:, :k])).to(gpu)
        capi.tune_check(dp, dk, k, 1, k, ok, force=20)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ok.cpu().numpy(), base[:k])
    # the timed tuner may keep the screen or not; the flags stay the same
    out.fill_(7)
    assert capi.tune_check(dp, dq, n, 1, n, out) in (1, 2, 4, 11, 12, 14, 20)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), base)
    if scene == "arm16":
        rng = np.random.default_rng(9)
        vox = torch.from_numpy(rng.uniform(-0.8, 0.8, (40, 3)).astype(np.float32)).to(gpu)
        rg = list(pl.getRobotGeometryIds())
        capi.tune_check(dp, dq, n, 1, n, out, force=20)
        out.fill_(7)
        capi.check_collision_free_voxels(dp, dq, n, 1, n, vox, 40, 0.03, rg, out)
        torch.cuda.synchronize()
        vnp = vox.cpu().numpy().T
        assert_flags_match(out.cpu().numpy(), orc.check_voxels(op, q, vnp, 0.03, rg),
                           orc.min_sd(op, q, vnp, 0.03, rg), "arm16 voxels after screening")
