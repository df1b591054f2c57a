"""The oracle's FP64 signed distance (orc.min_sd) pinned in magnitude by brute
force, for the three pair kinds the reference checks (collision_checker.cpp
:91-203): sphere-sphere, sphere-box and box-box.

min_sd sets the 1e-5 m contact band of every flag-parity test, so its value,
not only its sign, has to be right. Each case places two world-fixed
geometries on the obstacle link of a plant and makes them the plant's only
pair, so the oracle's minimum over pairs is their signed distance at any
configuration. The expected values never use the oracle's closed forms
(sqrt(d^2) - r, the per-axis clamp, the 15 SAT axes):

* a sphere of radius r and centre c against a convex body K has signed
  distance sd_point(c, K) - r; sd_point is +dist(c, dK) outside K and
  -dist(c, dK) inside, and dist(c, dK) is found by dense samples of the
  boundary dK (Fibonacci points on a sphere, a grid on each box face)
  refined by a bounded local search;
* two boxes: their signed distance is max over unit u of the gap
  min_{b in B} u.b - max_{a in A} u.a (the distance when separated, minus
  the penetration depth when overlapping). The support of a box is attained
  at a corner, so the gap is exact from its 8 corners; u runs over dense
  Fibonacci directions refined by a simplex search. When the boxes are
  separated the Euclidean distance also comes from a convex bounded
  minimisation over both boxes' local coordinates.

A mistake such as d^2 - r^2 in place of d - r fails the sphere cases by far
more than their tolerance."""
import numpy as np
import pytest
from scipy.optimize import minimize


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
    g.type = {"sphere": c.ShapePrimitive.SPHERE, "box": c.ShapePrimitive.BOX}[kind]
    g.dimensions = np.asarray(dims, np.float32)
    g.link_index = link
    g.group = c.CollisionGroup.WORLD
    g.X_L_B = np.asarray(X, np.float32)
    return g


def _oracle_pair(orc, base, a, b):
    pl, mp, geoms, _ = base
    mp.set_scene_collision_geometries(geoms + [a, b])
    n = len(geoms)
    mp.set_collision_pairs_from_matrix(np.array([[n], [n + 1]], np.int32))
    op = orc.Plant(mp)
    q = np.zeros((len(pl.getPositionLowerLimits()), 1), np.float32)
    return float(orc.min_sd(op, q)[0])


def _pose(rng, spread):
    from csdecomp_amd import scenes
    X = np.eye(4)
    X[:3, :3] = scenes.random_rotation(rng)
    X[:3, 3] = rng.uniform(-spread, spread, 3)
    return X.astype(np.float32).astype(np.float64)  # as the plant stores it


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def _fibonacci(n):
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = np.pi * (1.0 + 5.0 ** 0.5) * i
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


def _unit(a):
    """Unit vector from two angles (a search variable)."""
    t, p = a
    return np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])


def _angles(u):
    return np.array([np.arccos(np.clip(u[2], -1, 1)), np.arctan2(u[1], u[0])])


SPHERE_DIRS = _fibonacci(100_000)


def _dist_to_sphere_surface(c, centre, radius):
    """min over points of the sphere surface of |p - c| (samples + refine)."""
    pts = centre + radius * SPHERE_DIRS
    d = np.linalg.norm(pts - c, axis=1)
    k = int(np.argmin(d))
    f = lambda a: float(np.linalg.norm(centre + radius * _unit(a) - c))
    best = d[k]
    for start in (_angles(SPHERE_DIRS[k]),):
        r = minimize(f, start, method="Nelder-Mead",
                     options={"xatol": 1e-13, "fatol": 1e-15, "maxiter": 4000})
        best = min(best, r.fun)
    return best


FACE_GRID = np.linspace(-1.0, 1.0, 201)


def _dist_to_box_surface(c, X, half):
    """min over points of the box boundary of |p - c|: a 201 x 201 grid on
    each of the 6 faces, then a bounded search on the best few faces."""
    R, t = X[:3, :3], X[:3, 3]
    g0, g1 = np.meshgrid(FACE_GRID, FACE_GRID, indexing="ij")
    g0, g1 = g0.ravel(), g1.ravel()
    best = np.inf
    cands = []
    for axis in range(3):
        o = [k for k in range(3) if k != axis]
        for sgn in (-1.0, 1.0):
            loc = np.zeros((g0.size, 3))
            loc[:, axis] = sgn * half[axis]
            loc[:, o[0]] = g0 * half[o[0]]
            loc[:, o[1]] = g1 * half[o[1]]
            d = np.linalg.norm(loc @ R.T + t - c, axis=1)
            k = int(np.argmin(d))
            cands.append((d[k], axis, sgn, o, loc[k, o[0]], loc[k, o[1]]))
            best = min(best, d[k])
    cands.sort(key=lambda x: x[0])
    for _, axis, sgn, o, s0, s1 in cands[:3]:
        def f(x, axis=axis, sgn=sgn, o=o):
            loc = np.zeros(3)
            loc[axis] = sgn * half[axis]
            loc[o[0]], loc[o[1]] = x
            return float(np.sum((R @ loc + t - c) ** 2))
        r = minimize(f, np.array([s0, s1]), method="L-BFGS-B",
                     bounds=[(-half[o[0]], half[o[0]]), (-half[o[1]], half[o[1]])],
                     options={"ftol": 1e-16, "gtol": 1e-13, "maxiter": 2000})
        best = min(best, float(np.sqrt(max(r.fun, 0.0))))
    return best


def _box_corners(X, half):
    s = np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1) for k in (-1, 1)], float)
    return (s * half) @ X[:3, :3].T + X[:3, 3]


BOX_DIRS = _fibonacci(60_000)


def _box_box_sd_brute(XA, hA, XB, hB):
    """max over unit u of min_{b in B} u.b - max_{a in A} u.a, from the
    corners; returns (value, maximising direction)."""
    CA, CB = _box_corners(XA, hA), _box_corners(XB, hB)
    gap = (BOX_DIRS @ CB.T).min(axis=1) - (BOX_DIRS @ CA.T).max(axis=1)
    f = lambda a: -float((CB @ _unit(a)).min() - (CA @ _unit(a)).max())
    best, best_u = -np.inf, None
    for k in np.argsort(-gap)[:6]:
        r = minimize(f, _angles(BOX_DIRS[k]), method="Nelder-Mead",
                     options={"xatol": 1e-14, "fatol": 1e-16, "maxiter": 6000})
        for val, u in ((-r.fun, _unit(r.x)), (gap[k], BOX_DIRS[k])):
            if val > best:
                best, best_u = val, u
    return best, best_u


def _box_box_distance_qp(XA, hA, XB, hB):
    """Euclidean distance of two separated boxes: a convex bounded problem
    over both boxes' local coordinates."""
    RA, tA, RB, tB = XA[:3, :3], XA[:3, 3], XB[:3, :3], XB[:3, 3]
    f = lambda x: float(np.sum((RA @ x[:3] + tA - RB @ x[3:] - tB) ** 2))
    best = np.inf
    for x0 in (np.zeros(6), np.concatenate([RA.T @ (tB - tA), RB.T @ (tA - tB)])):
        x0 = np.clip(x0, -np.concatenate([hA, hB]), np.concatenate([hA, hB]))
        r = minimize(f, x0, method="L-BFGS-B",
                     bounds=[(-h, h) for h in np.concatenate([hA, hB])],
                     options={"ftol": 1e-18, "gtol": 1e-14, "maxiter": 5000})
        best = min(best, float(np.sqrt(max(r.fun, 0.0))))
    return best


def test_sphere_sphere_magnitude(orc, base):
    rng = np.random.default_rng(41)
    link = base[3]
    seen = [0, 0]
    for case in range(120):
        XA, XB = _pose(rng, 0.25), _pose(rng, 0.25)
        rA, rB = _f32(rng.uniform(0.03, 0.3)), _f32(rng.uniform(0.03, 0.3))
        sd = _oracle_pair(orc, base, _geom("sphere", (rA, -1, -1), XA, link),
                          _geom("sphere", (rB, -1, -1), XB, link))
        cA, cB = XA[:3, 3], XB[:3, 3]
        d = _dist_to_sphere_surface(cA, cB, rB)
        inside = np.linalg.norm(cA - cB) < rB
        ref = (-d if inside else d) - rA
        assert abs(sd - ref) < 1e-7, (case, sd, ref)
        seen[int(ref > 0)] += 1
    assert min(seen) >= 20  # both overlapping and separated pairs ran


def test_sphere_box_magnitude(orc, base):
    rng = np.random.default_rng(42)
    link = base[3]
    kinds = {"separated": 0, "overlap_outside": 0, "centre_inside": 0}
    for case in range(150):
        XB = _pose(rng, 0.2)
        half = _f32(2 * rng.uniform(0.05, 0.3, 3)) / 2
        if case % 3 == 0:  # centre inside the box
            loc = rng.uniform(-0.95, 0.95, 3) * half
        else:
            loc = rng.uniform(-1.6, 1.6, 3) * (half + 0.1)
        XS = np.eye(4)
        XS[:3, 3] = XB[:3, :3] @ loc + XB[:3, 3]
        XS = XS.astype(np.float32).astype(np.float64)
        r = _f32(rng.uniform(0.02, 0.2))
        sd = _oracle_pair(orc, base, _geom("sphere", (r, -1, -1), XS, link),
                          _geom("box", 2 * half, XB, link))
        c = XS[:3, 3]
        d = _dist_to_box_surface(c, XB, half)
        l = XB[:3, :3].T @ (c - XB[:3, 3])
        inside = bool(np.all(np.abs(l) <= half))
        ref = (-d if inside else d) - r
        assert abs(sd - ref) < 1e-7, (case, sd, ref, inside)
        kinds["centre_inside" if inside else ("separated" if ref > 0 else "overlap_outside")] += 1
    assert min(kinds.values()) >= 15, kinds


def test_box_box_sat_semantics(orc, base):
    """Overlapping boxes: the oracle's value is minus the penetration depth,
    exactly (the minimum translation of two convex polyhedra is along a face
    normal or an edge-edge cross product, all among the 15 SAT axes).
    Separated boxes: the oracle's value is positive and a lower bound of the
    Euclidean distance (the 15 axes need not hold the best separating
    direction), and equal to it when the best direction is a face normal."""
    rng = np.random.default_rng(43)
    link = base[3]
    seen = {"overlap": 0, "separated": 0, "face_normal": 0}
    for case in range(120):
        XA, XB = _pose(rng, 0.35), _pose(rng, 0.35)
        hA = _f32(2 * rng.uniform(0.05, 0.3, 3)) / 2
        hB = _f32(2 * rng.uniform(0.05, 0.3, 3)) / 2
        sd = _oracle_pair(orc, base, _geom("box", 2 * hA, XA, link), _geom("box", 2 * hB, XB, link))
        ref, u = _box_box_sd_brute(XA, hA, XB, hB)
        if ref < -1e-6:
            assert abs(sd - ref) < 1e-7, (case, sd, ref)
            seen["overlap"] += 1
        elif ref > 1e-6:
            dist = _box_box_distance_qp(XA, hA, XB, hB)
            assert abs(dist - ref) < 1e-6, (case, dist, ref)  # the two brute forces agree
            assert 0.0 < sd <= dist + 1e-9, (case, sd, dist)
            normals = np.concatenate([XA[:3, :3].T, XB[:3, :3].T])
            if np.max(np.abs(normals @ u)) > 1 - 1e-9:
                assert abs(sd - dist) < 1e-6, (case, sd, dist)
                seen["face_normal"] += 1
            seen["separated"] += 1
    assert seen["overlap"] >= 20 and seen["separated"] >= 20 and seen["face_normal"] >= 5, seen

