"""GPU parity of the voxel collision maps (csd_voxel_collision_map,
RoadmapBuilder.build_collision_map / build_edge_collision_map) against the
oracle, which checks every voxel at every configuration; entries may differ
only where the voxel sphere is within 1e-5 m of contact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CENTER, SIZE, RES = np.array([0.0, 0.0, 0.4], np.float32), [1.6, 1.6, 1.6], 0.05


def _gpu_map(capi, gpu, mp, rg, nodes, edges, step):
    import torch
    nx = int(np.ceil(np.float32(1.6) / np.float32(0.1)))
    V = nx ** 3
    dn = torch.from_numpy(np.ascontiguousarray(nodes, np.float32)).to(gpu)
    de = None if edges is None else torch.from_numpy(np.ascontiguousarray(edges, np.int32)).to(gpu)
    off = torch.empty(V + 1, dtype=torch.int64, device=gpu)
    dp = capi.DevicePlant(mp)
    ne = 0 if edges is None else len(edges)
    total = capi.voxel_collision_map(dp, dn, len(nodes), de, ne, step, rg, CENTER, SIZE, RES, off)
    items = torch.empty(max(total, 1), dtype=torch.int32, device=gpu)
    if total > 1:
        with pytest.raises(capi.CsdError):  # too small: the count is still reported
            capi.voxel_collision_map(dp, dn, len(nodes), de, ne, step, rg, CENTER, SIZE, RES, off,
                                     items=items, capacity=total - 1)
    assert capi.voxel_collision_map(dp, dn, len(nodes), de, ne, step, rg, CENTER, SIZE, RES, off,
                                    items=items, capacity=total) == total
    o, it = off.cpu().numpy(), items.cpu().numpy()
    return {int(v): sorted(it[o[v]:o[v + 1]].tolist()) for v in np.nonzero(np.diff(o))[0]}


def _assert_maps_match(orc, op, got, ref, configs, rg):
    _, centres = orc.roadmap_grid(CENTER, SIZE, RES)
    radius = float(np.float32(np.sqrt(3.0) * np.float32(RES)))
    diff = 0
    for v in set(got) | set(ref):
        for item in set(got.get(v, [])) ^ set(ref.get(v, [])):
            q = configs(item)
            lo = orc.voxel_flags(op, q, centres[:, v:v + 1], radius - 1e-5, rg)
            hi = orc.voxel_flags(op, q, centres[:, v:v + 1], radius + 1e-5, rg)
            assert (lo != hi).any(), f"voxel {v} item {item} differs outside the contact band"
            diff += 1
    assert diff <= 3


def _scene():
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    return pl, pl.getMinimalPlant(), list(pl.getSceneInspector().robot_geometry_ids)


def test_node_map_parity(gpu, orc):
    from csdecomp_amd import capi, scenes
    pl, mp, rg = _scene()
    op = orc.Plant(mp)
    q = scenes.uniform_configs(31, 300, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    nodes = np.ascontiguousarray(q.T)
    got = _gpu_map(capi, gpu, mp, rg, nodes, None, 1.0)
    ref = orc.voxel_map(op, nodes, None, 1.0, rg, CENTER, SIZE, RES)
    assert sum(len(v) for v in ref.values()) > 1000
    _assert_maps_match(orc, op, got, ref, lambda i: nodes[i][:, None], rg)


def test_edge_map_parity(gpu, orc):
    from csdecomp_amd import capi, scenes
    pl, mp, rg = _scene()
    op = orc.Plant(mp)
    rng = np.random.default_rng(32)
    lo = np.float32(pl.getPositionLowerLimits())
    hi = np.float32(pl.getPositionUpperLimits())
    a = rng.uniform(lo, hi, (40, 7)).astype(np.float32)
    d = rng.normal(size=(40, 7))
    d *= rng.uniform(0.02, 0.5, (40, 1)) / np.linalg.norm(d, axis=1, keepdims=True)
    nodes = np.concatenate([a, np.clip(a + d, lo, hi).astype(np.float32)])
    edges = np.array([(i, i + 40) for i in range(40)] + [(0, 0), (3, 44)], np.int32)
    step = 0.02
    got = _gpu_map(capi, gpu, mp, rg, nodes, edges, step)
    ref = orc.voxel_map(op, nodes, edges, step, rg, CENTER, SIZE, RES)
    assert sum(len(v) for v in ref.values()) > 500
    _assert_maps_match(orc, op, got, ref,
                       lambda e: orc.roadmap_edge_configs(nodes[edges[e][0]], nodes[edges[e][1]],
                                                          step), rg)


def test_roadmap_builder_flow(gpu, orc, csd):
    """add_nodes_manual, build_roadmap (against the oracle's BuildRoadmap),
    the two voxel maps with their drop rules, and the pose map."""
    from csdecomp_amd import scenes
    pl, mp, rg = _scene()
    opts = csd.RoadmapOptions()
    opts.map_center = CENTER
    opts.robot_map_size_x = opts.robot_map_size_y = opts.robot_map_size_z = 1.6
    opts.offline_voxel_resolution = RES
    opts.max_configuration_distance_between_nodes = 4.0
    opts.max_task_space_distance_between_nodes = 0.6
    opts.edge_step_size = 0.05
    b = csd.RoadmapBuilder(pl, pl.getKinematicTree().get_link_name(7), opts)
    q = scenes.uniform_configs(33, 20000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    added = b.add_nodes_manual(q)
    free = np.asarray(csd.CheckCollisionFreeCuda(q, mp), bool)
    assert added == free.sum() > 100
    X = np.ascontiguousarray(q[:, free].T)
    b.build_roadmap(5)
    drm = b.get_drm()
    op = orc.Plant(mp)
    ref = orc.build_roadmap(op, X, 5, 4.0, 0.6, 0.05)
    ids = sorted(drm.id_to_node_map)
    got = {k - 1: sorted(j - 1 for j in v) for k, v in drm.node_adjacency_map.items()}
    assert {k: sorted(v) for k, v in ref.items()} == got
    assert len(got) > 20 and all(len(v) >= 1 for v in got.values())
    b.build_collision_map()
    nodes = np.array([drm.id_to_node_map[i] for i in ids])
    cm = orc.voxel_map(op, nodes, None, 1.0, rg, CENTER, SIZE, RES)
    expect = {v: [ids[i] for i in lst] for v, lst in cm.items() if len(lst) < len(ids)}
    assert len(set(drm.collision_map) ^ set(expect)) <= 3
    assert sum(drm.collision_map.get(v) == expect.get(v) for v in expect) >= len(expect) - 3
    b.build_edge_collision_map(0.05)
    ecm = drm.edge_collision_map
    assert ecm and all((b_, a_) in v for v in ecm.values() for a_, b_ in v)
    b.build_pose_map()
    t = drm.target_link_index_
    for k in (0, len(ids) - 1):
        tf = csd.computeLinkFrameToWorldTransformsMinimal(nodes[k], mp.kin_tree)
        np.testing.assert_allclose(drm.id_to_pose_map[ids[k]], np.asarray(tf)[4 * t:4 * t + 4],
                                   atol=1e-5)


def test_knn_pages(gpu, orc):
    """Successive csd_knn_page calls walk each node's full sorted list."""
    import torch
    from csdecomp_amd import capi
    rng = np.random.default_rng(41)
    X = rng.uniform(-1, 1, (500, 7)).astype(np.float32)
    X[7] = X[3]  # an exact duplicate is skipped (d^2 <= FLT_EPSILON)
    dX = torch.from_numpy(X).to(gpu)
    queries = np.array([0, 3, 250, 499], np.int32)
    got = {int(qn): [] for qn in queries}
    ad = ai = None
    for _ in range(4):
        oi = torch.empty(len(queries) * 16, dtype=torch.int32, device=gpu)
        od = torch.empty(len(queries) * 16, dtype=torch.float32, device=gpu)
        capi.knn_page(dX, 500, 7, torch.from_numpy(queries).to(gpu), len(queries), 16,
                      None if ad is None else torch.from_numpy(ad).to(gpu),
                      None if ai is None else torch.from_numpy(ai).to(gpu), oi, od)
        idx = oi.cpu().numpy().reshape(-1, 16)
        d2 = od.cpu().numpy().reshape(-1, 16)
        for r, qn in enumerate(queries):
            got[int(qn)] += idx[r].tolist()
        ad, ai = d2[:, -1].copy(), idx[:, -1].copy()
    for qn in queries:
        d2 = np.zeros(500, np.float32)
        for d in range(7):
            e = (X[:, d] - X[qn, d]).astype(np.float32)
            d2 = (d2 + e * e).astype(np.float32)
        order = [j for j in np.lexsort((np.arange(500), d2)) if d2[j] > np.float32(1.1920928955078125e-07)]
        assert got[int(qn)] == order[:64], qn


def test_random_samples_and_archive(gpu, orc, csd, tmp_path):
    """get_random_samples (roadmap_builder.cpp:53-69): nodes 1..n are the
    collision-free hit-and-run samples, inside the joint limits and free per
    the oracle's exact check; the DRM then survives a Write/Read round trip."""
    pl, mp, rg = _scene()
    opts = csd.RoadmapOptions()
    opts.max_configuration_distance_between_nodes = 4.0
    opts.max_task_space_distance_between_nodes = 0.6
    b = csd.RoadmapBuilder(pl, pl.getKinematicTree().get_link_name(7), opts)
    b.get_random_samples(8000)
    raw = dict(b._raw)
    ids = sorted(raw)
    assert 40 < len(ids) < 8000 and ids == list(range(1, len(ids) + 1))
    X = np.array([raw[i] for i in ids], np.float32)
    lo = np.asarray(pl.getPositionLowerLimits(), np.float32)
    hi = np.asarray(pl.getPositionUpperLimits(), np.float32)
    assert np.all((X >= lo - 1e-6) & (X <= hi + 1e-6))
    op = orc.Plant(mp)
    q = np.asfortranarray(X.T)
    ref = orc.check(op, q, orc.RULE_GPU)
    near = np.abs(orc.min_sd(op, q)) < 1e-5
    assert np.all(ref.astype(bool) | near)
    b.build_roadmap(4)
    b.build_pose_map()
    drm = b.get_drm()
    ids = sorted(drm.id_to_node_map)  # only nodes the greedy connected (roadmap_builder.cpp:278)
    assert ids and set(ids) <= set(raw)
    assert all(np.array_equal(drm.id_to_node_map[i], raw[i]) for i in ids)
    p = tmp_path / "r.drm"
    b.write(str(p))
    e = csd.DRM()
    e.Read(str(p))
    assert e.node_adjacency_map == {k: list(v) for k, v in drm.node_adjacency_map.items()}
    for i in (ids[0], ids[-1]):
        np.testing.assert_array_equal(e.id_to_node_map[i], drm.id_to_node_map[i])
        np.testing.assert_array_equal(e.id_to_pose_map[i], drm.id_to_pose_map[i])
    assert e.target_link_index_ == drm.target_link_index_
