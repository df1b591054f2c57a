"""Seeded synthetic scenes of the benchmark workloads, built through the
same plant API a user would call (KinematicTree + finalize_scene + Plant).

* ``arm_scene``: 7-DOF serial revolute arm, link lengths U[0.1, 0.4] m, joint
  axes alternating z/y, limits [-pi, pi], 4 spheres per link (base
  included; radius U[0.03, 0.08] m, centres along the link), plus world-fixed
  oriented boxes (half extents U[0.05, 0.3] m, centres U[-1, 1]^3 m outside
  a 0.2 m ball, uniform random rotation) and spheres (radius U[0.05, 0.2] m).
* ``dual_arm_scene``: two such arms with bases at (+-0.3, 0, 0) and a
  sparser obstacle field (see its docstring).
* ``boxes_in_corners``: the 2-DOF scene of the reference's EI-ZO tests
  (planning/tests/edge_inflation_zero_order_test.cpp:72-115).
"""
import math

import numpy as np

from . import _pycsdecomp as _c


def _joint(name, kind, parent, child, xyz=(0, 0, 0), axis=(0, 0, 1), lo=-math.pi, hi=math.pi):
    j = _c.Joint()
    j.name = name
    j.type = kind
    X = np.eye(4, dtype=np.float32)
    X[:3, 3] = xyz
    j.X_PL_J = X
    j.parent_link = parent
    j.child_link = child
    j.axis = np.asarray(axis, np.float32)
    j.position_lower_limit = lo
    j.position_upper_limit = hi
    return j


def _link(tree, name):
    l = _c.Link()
    l.name = name
    l.parent_joint = -1
    tree.add_link(l)
    return tree.get_link_index(name)


def _sphere(link, center, r):
    g = _c.CollisionGeometry()
    g.type = _c.ShapePrimitive.SPHERE
    g.dimensions = np.array([r, -1, -1], np.float32)
    g.link_index = link
    X = np.eye(4, dtype=np.float32)
    X[:3, 3] = center
    g.X_L_B = X
    return g


def _box(link, center, half, R):
    g = _c.CollisionGeometry()
    g.type = _c.ShapePrimitive.BOX
    g.dimensions = (2.0 * np.asarray(half)).astype(np.float32)
    g.link_index = link
    X = np.eye(4)
    X[:3, :3] = R
    X[:3, 3] = center
    g.X_L_B = X.astype(np.float32)
    return g


def random_rotation(rng):
    """Uniform random rotation from a uniform quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.uniform(size=3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    x, y, z, w = (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                  b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _add_arm(tree, geoms, rng, prefix, base_xyz, parent=0):
    """world -fixed-> base -rev z-> l1 -rev y-> l2 ... (7 revolute joints)."""
    base = _link(tree, f"{prefix}base")
    tree.add_joint(_joint(f"{prefix}base_weld", _c.JointType.FIXED, parent, base, base_xyz))
    lengths = rng.uniform(0.1, 0.4, size=8)  # base + 7 links
    for k in range(4):
        geoms.append(_sphere(base, (0, 0, rng.uniform(0, lengths[0])), rng.uniform(0.03, 0.08)))
    prev, offset = base, (0.0, 0.0, lengths[0])
    for i in range(7):
        lk = _link(tree, f"{prefix}l{i + 1}")
        axis = (0, 0, 1) if i % 2 == 0 else (0, 1, 0)
        tree.add_joint(_joint(f"{prefix}j{i + 1}", _c.JointType.REVOLUTE, prev, lk, offset, axis))
        ln = lengths[i + 1]
        for k in range(4):
            geoms.append(_sphere(lk, (rng.uniform(0, ln), 0, 0), rng.uniform(0.03, 0.08)))
        prev, offset = lk, (ln, 0.0, 0.0)


def _add_obstacles(tree, geoms, rng, num_boxes, num_spheres, half=(0.05, 0.3),
                   radius=(0.05, 0.2), extent=1.0, keep_out=((0.0, 0.0, 0.0),), keep_r=0.2):
    """World-fixed boxes (half extents U[half], uniform random rotation) and
    spheres (radius U[radius]); centres uniform in [-extent, extent]^3 at
    least keep_r away from every point of keep_out (the arm bases)."""
    obs = _link(tree, "obstacles")
    tree.add_joint(_joint("obstacle_weld", _c.JointType.FIXED, 0, obs))
    ko = np.asarray(keep_out, np.float64)

    def centre():
        while True:
            c = rng.uniform(-extent, extent, size=3)
            if np.linalg.norm(ko - c, axis=1).min() >= keep_r:
                return c

    for _ in range(num_boxes):
        c = centre()
        geoms.append(_box(obs, c, rng.uniform(half[0], half[1], size=3), random_rotation(rng)))
    for _ in range(num_spheres):
        c = centre()
        geoms.append(_sphere(obs, c, rng.uniform(radius[0], radius[1])))


def _finish(tree, geoms):
    tree.finalize()
    names = {f"geom{i}": i for i in range(len(geoms))}
    geoms, insp = _c.finalize_scene(tree, geoms, names)
    return _c.Plant(tree, insp, list(geoms))


def arm_scene(seed, num_boxes=16, num_spheres=0):
    """Single 7-DOF arm + world-fixed obstacles; returns a Plant."""
    rng = np.random.default_rng(seed)
    tree = _c.KinematicTree()
    _link(tree, "world")
    geoms = []
    _add_arm(tree, geoms, rng, "", (0, 0, 0))
    _add_obstacles(tree, geoms, rng, num_boxes, num_spheres)
    return _finish(tree, geoms)


def dual_arm_scene(seed, num_boxes=72, num_spheres=24, half=(0.05, 0.2), radius=(0.05, 0.15),
                   extent=1.5, keep_r=0.4):
    """Two 7-DOF arms with bases 0.6 m apart + obstacles; returns a Plant.

    With the single-arm obstacle distribution (half extents U[0.05, 0.3] m,
    radius U[0.05, 0.2] m, centres in [-1, 1]^3), 96 obstacles leave no free
    configuration for two arms (none in 100,000 uniform samples). The
    defaults here (smaller obstacles over [-1.5, 1.5]^3, centres at least
    0.4 m from either base) leave about 4% of configurations free, close to
    the single-arm benchmark scene."""
    rng = np.random.default_rng(seed)
    tree = _c.KinematicTree()
    _link(tree, "world")
    geoms = []
    _add_arm(tree, geoms, rng, "a_", (0.3, 0, 0))
    _add_arm(tree, geoms, rng, "b_", (-0.3, 0, 0))
    _add_obstacles(tree, geoms, rng, num_boxes, num_spheres, half, radius, extent,
                   ((0.3, 0.0, 0.0), (-0.3, 0.0, 0.0)), keep_r)
    return _finish(tree, geoms)


BOXES_IN_CORNERS_URDF = """
<robot name="boxes">
  <link name="fixed">
    <collision name="top_left"><origin rpy="0 0 0" xyz="-1 1 0"/><geometry><box size="1.4 1.4 1.4"/></geometry></collision>
    <collision name="top_right"><origin rpy="0 0 0" xyz="1 1 0"/><geometry><box size="1.4 1.4 1.4"/></geometry></collision>
    <collision name="bottom_left"><origin rpy="0 0 0" xyz="-1 -1 0"/><geometry><box size="1.4 1.4 1.4"/></geometry></collision>
    <collision name="bottom_right"><origin rpy="0 0 0" xyz="1 -1 0"/><geometry><box size="1.4 1.4 1.4"/></geometry></collision>
  </link>
  <joint name="fixed_link_weld" type="fixed"><parent link="world"/><child link="fixed"/></joint>
  <link name="movable"><collision name="sphere"><geometry><sphere radius="0.01"/></geometry></collision></link>
  <link name="for_joint"/>
  <joint name="x" type="prismatic"><axis xyz="1 0 0"/><limit lower="-2" upper="2"/><parent link="world"/><child link="for_joint"/></joint>
  <joint name="y" type="prismatic"><axis xyz="0 1 0"/><limit lower="-2" upper="2"/><parent link="for_joint"/><child link="movable"/></joint>
</robot>
"""


def boxes_in_corners():
    """URDFParser of the reference's 2-D EI-ZO test scene."""
    from .urdf import URDFParser
    p = URDFParser()
    p.parse_urdf_string(BOXES_IN_CORNERS_URDF)
    return p


def uniform_configs(seed, n, lower, upper):
    """(dof, n) float32 configurations uniform in the box."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lower, np.float64), np.asarray(upper, np.float64)
    return np.asfortranarray(rng.uniform(lo[:, None], hi[:, None], size=(len(lo), n)).astype(np.float32))


def free_segments(plant, count, seed, length=(0.2, 1.5), log_uniform=False, step=0.01,
                  batch=1 << 18):
    """`count` straight joint-space segments whose endpoints are collision
    free and whose checkEdgeCollisionFree discretisation at `step` is free,
    checked on the GPU. Starts are uniform in the joint limits; directions
    are uniform on the sphere; lengths are U[length] (log-uniform when
    `log_uniform`); ends are clipped to the limits. Returns two (dim, count)
    float32 arrays (starts, ends)."""
    import torch
    from . import capi

    mp = plant.getMinimalPlant()
    lo = np.asarray(plant.getPositionLowerLimits(), np.float32)
    hi = np.asarray(plant.getPositionUpperLimits(), np.float32)
    dim = len(lo)
    rng = np.random.default_rng(seed)
    dp = capi.DevicePlant(mp)
    dev = torch.device("cuda", torch.cuda.current_device())
    starts, ends = [], []
    while len(starts) < count:
        q1 = rng.uniform(lo, hi, size=(batch, dim)).astype(np.float32)
        d = rng.normal(size=(batch, dim))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        if log_uniform:
            ln = np.exp(rng.uniform(np.log(length[0]), np.log(length[1]), size=(batch, 1)))
        else:
            ln = rng.uniform(length[0], length[1], size=(batch, 1))
        q2 = np.clip(q1 + ln * d, lo, hi).astype(np.float32)
        nodes = torch.from_numpy(np.concatenate([q1, q2])).to(dev)
        free = torch.empty(2 * batch, dtype=torch.uint8, device=dev)
        capi.check_collision_free(dp, nodes, 2 * batch, dim, 1, free)
        ok = (free[:batch] & free[batch:]).nonzero().flatten().to(torch.int32)
        if ok.numel() == 0:
            continue
        edges = torch.stack([ok, ok + batch], dim=1).contiguous()
        valid = torch.empty(ok.numel(), dtype=torch.uint8, device=dev)
        capi.validate_edges(dp, nodes, 2 * batch, edges, ok.numel(), step, 0, valid)
        idx = ok[valid.bool()].cpu().numpy()
        for i in idx[: count - len(starts)]:
            starts.append(q1[i])
            ends.append(q2[i])
    return (np.asfortranarray(np.array(starts, np.float32).T),
            np.asfortranarray(np.array(ends, np.float32).T))


def log_uniform_budgets(seed, count, low=512, high=16384):
    """Per-segment particle budgets, log-uniform integers in [low, high]."""
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(x))) for x in rng.uniform(math.log(low), math.log(high), count)]
