"""The oracle's capsule distances pinned by brute force. csdecomp has no
capsule code (its CPU checker throws on them, collision_checker.cpp:104-107),
so nothing of the reference can pin them.

Each case places two world-fixed geometries on the obstacle link of a plant
and makes them its only pair, so the oracle's minimum signed distance over
pairs is their distance at any configuration. The expected value comes from
dense samples along the capsule segments (refined by a bounded search for
capsule-capsule), never from the oracle's closed forms: point-segment
projection (capsule-sphere), the clamped segment-segment solution
(capsule-capsule) and the golden-section search of segment-box."""
import numpy as np
import pytest
from scipy.optimize import minimize, minimize_scalar


@pytest.fixture(scope="module")
def base():
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1000, 1, 0)
    mp = pl.getMinimalPlant()
    geoms = list(mp.scene_collision_geometries)
    return pl, mp, geoms, geoms[-1].link_index


def _geom(kind, dims, X, link):
    from csdecomp_amd import _pycsdecomp as c
    g = c.CollisionGeometry()
    g.type = {"capsule": c.ShapePrimitive.CAPSULE, "sphere": c.ShapePrimitive.SPHERE,
              "box": c.ShapePrimitive.BOX}[kind]
    g.dimensions = np.asarray(dims, np.float32)
    g.link_index = link
    g.group = c.CollisionGroup.WORLD
    g.X_L_B = np.asarray(X, np.float32)
    return g


def _oracle_pair(orc, base, a, b):
    """(signed distance in FP64, GPU-rule flag) of the pair (a, b) alone."""
    pl, mp, geoms, _ = base
    mp.set_scene_collision_geometries(geoms + [a, b])
    n = len(geoms)
    mp.set_collision_pairs_from_matrix(np.array([[n], [n + 1]], np.int32))
    op = orc.Plant(mp)
    q = np.zeros((len(pl.getPositionLowerLimits()), 1), np.float32)
    return float(orc.min_sd(op, q)[0]), int(orc.check(op, q, orc.RULE_GPU)[0])


def _pose(rng, spread=0.4):
    from csdecomp_amd import scenes
    X = np.eye(4)
    X[:3, :3] = scenes.random_rotation(rng)
    X[:3, 3] = rng.uniform(-spread, spread, 3)
    return X.astype(np.float32).astype(np.float64)


def _f32(*v):
    """Values as the plant stores them (float32)."""
    return [float(np.float32(x)) for x in v]


def _ends(X, length):
    c, z = X[:3, 3], X[:3, 2]
    return c - 0.5 * length * z, c + 0.5 * length * z


def _seg(p0, p1, t):
    return p0[None, :] + np.asarray(t)[:, None] * (p1 - p0)[None, :]


def _box_sd(X, half, pts):
    """Signed distance of points to an oriented box (negative inside)."""
    loc = (pts - X[:3, 3]) @ X[:3, :3]
    e = np.abs(loc) - half
    outside = np.linalg.norm(np.maximum(e, 0.0), axis=1)
    inside = np.minimum(e.max(axis=1), 0.0)
    return outside + inside


def _capsule_point_brute(p0, p1, c):
    t = np.linspace(0.0, 1.0, 200001)
    d = np.linalg.norm(_seg(p0, p1, t) - c, axis=1)
    k = int(np.argmin(d))
    lo, hi = t[max(k - 1, 0)], t[min(k + 1, len(t) - 1)]
    f = lambda s: float(np.linalg.norm(p0 + s * (p1 - p0) - c))
    r = minimize_scalar(f, bounds=(lo, hi), method="bounded", options={"xatol": 1e-13})
    return min(d[k], r.fun)


def _capsule_capsule_brute(a0, a1, b0, b1):
    t = np.linspace(0.0, 1.0, 801)
    A, B = _seg(a0, a1, t), _seg(b0, b1, t)
    D = np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2)
    i, j = np.unravel_index(int(np.argmin(D)), D.shape)
    f = lambda x: float(np.linalg.norm(a0 + x[0] * (a1 - a0) - b0 - x[1] * (b1 - b0)))
    r = minimize(f, np.array([t[i], t[j]]), method="L-BFGS-B", bounds=[(0, 1), (0, 1)],
                 options={"ftol": 1e-15, "gtol": 1e-12})
    return min(D[i, j], r.fun)


def test_capsule_sphere_brute_force(orc, base):
    rng = np.random.default_rng(31)
    link = base[3]
    for case in range(40):
        XA = _pose(rng)
        L = 0.0 if case == 0 else rng.uniform(0.05, 0.6)  # case 0: a zero-length capsule
        L, rA, rB = _f32(L, rng.uniform(0.02, 0.15), rng.uniform(0.02, 0.15))
        XB = _pose(rng)
        a = _geom("capsule", (rA, L, rA), XA, link)
        b = _geom("sphere", (rB, -1, -1), XB, link)
        sd, flag = _oracle_pair(orc, base, a, b)
        p0, p1 = _ends(XA, L)
        ref = _capsule_point_brute(p0, p1, XB[:3, 3]) - rA - rB
        assert abs(sd - ref) < 1e-7, (case, sd, ref)
        if abs(ref) > 1e-4:
            assert flag == int(ref > 0), (case, ref, flag)


def test_capsule_capsule_brute_force(orc, base):
    rng = np.random.default_rng(32)
    link = base[3]
    for case in range(30):
        XA = _pose(rng)
        XB = _pose(rng)
        if case < 4:  # parallel axes: the degenerate case of the closed form
            XB[:3, :3] = XA[:3, :3]
        LA, LB, rA, rB = _f32(rng.uniform(0.05, 0.6), rng.uniform(0.05, 0.6),
                              rng.uniform(0.02, 0.12), rng.uniform(0.02, 0.12))
        a = _geom("capsule", (rA, LA, rA), XA, link)
        b = _geom("capsule", (rB, LB, rB), XB, link)
        sd, flag = _oracle_pair(orc, base, a, b)
        a0, a1 = _ends(XA, LA)
        b0, b1 = _ends(XB, LB)
        ref = _capsule_capsule_brute(a0, a1, b0, b1) - rA - rB
        assert abs(sd - ref) < 1e-7, (case, sd, ref)
        if abs(ref) > 1e-4:
            assert flag == int(ref > 0), (case, ref, flag)


def test_capsule_box_brute_force(orc, base):
    rng = np.random.default_rng(33)
    link = base[3]
    t = np.linspace(0.0, 1.0, 200001)
    seen_inside = seen_outside = 0
    for case in range(40):
        XA = _pose(rng)
        XB = _pose(rng, 0.25)
        L, r = _f32(rng.uniform(0.05, 0.6), rng.uniform(0.02, 0.12))
        half = rng.uniform(0.05, 0.3, 3)
        a = _geom("capsule", (r, L, r), XA, link)
        b = _geom("box", 2 * half, XB, link)
        sd, flag = _oracle_pair(orc, base, a, b)
        p0, p1 = _ends(XA, L)
        # float32 box extents as the plant stores them
        h = np.asarray(2 * half, np.float32).astype(np.float64) / 2
        ref = float(_box_sd(XB, h, _seg(p0, p1, t)).min()) - r
        # the sample spacing bounds the brute force's excess (the signed
        # distance is 1-Lipschitz along the segment)
        assert ref - L / 200000 - 1e-9 <= sd <= ref + 1e-9, (case, sd, ref)
        if abs(ref) > 1e-4:
            assert flag == int(ref > 0), (case, ref, flag)
        seen_inside += ref < 0
        seen_outside += ref > 0
    assert seen_inside and seen_outside  # both penetrating and separated cases ran
