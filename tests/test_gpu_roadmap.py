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
    from csdecomp_amd import scenes
    pl, mp, rg = _scene()
    opts = csd.RoadmapOptions()
    opts.map_center = CENTER
    opts.robot_map_size_x = opts.robot_map_size_y = opts.robot_map_size_z = 1.6
    opts.offline_voxel_resolution = RES
    b = csd.RoadmapBuilder(pl, pl.getKinematicTree().get_link_name(7), opts)
    q = scenes.uniform_configs(33, 4000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    added = b.add_nodes_manual(q)
    drm = b.get_drm()
    free = np.asarray(csd.CheckCollisionFreeCuda(q, mp), bool)
    assert added == free.sum() == len(drm.id_to_node_map) > 10
    ids = sorted(drm.id_to_node_map)
    X = np.array([drm.id_to_node_map[i] for i in ids])
    # a ring roadmap: each node to its successor
    drm.node_adjacency_map = {i: [] for i in ids}
    for k in range(len(ids)):
        a, c = ids[k], ids[(k + 1) % len(ids)]
        drm.node_adjacency_map[a].append(c)
        drm.node_adjacency_map[c].append(a)
    b.build_collision_map()
    op = orc.Plant(mp)
    ref = orc.voxel_map(op, X, None, 1.0, rg, CENTER, SIZE, RES)
    expect = {v: [ids[i] for i in lst] for v, lst in ref.items() if len(lst) < len(ids)}
    assert len(set(drm.collision_map) ^ set(expect)) <= 3
    agree = sum(drm.collision_map.get(v) == expect.get(v) for v in expect)
    assert agree >= len(expect) - 3
    b.build_edge_collision_map(0.05)
    ecm = drm.edge_collision_map
    assert ecm and all((b_, a_) in v for v in ecm.values() for a_, b_ in v)
    b.build_pose_map()
    t = drm.target_link_index_
    for k in (0, len(ids) - 1):
        tf = csd.computeLinkFrameToWorldTransformsMinimal(X[k], mp.kin_tree)
        np.testing.assert_allclose(drm.id_to_pose_map[ids[k]], np.asarray(tf)[4 * t:4 * t + 4],
                                   atol=1e-5)
