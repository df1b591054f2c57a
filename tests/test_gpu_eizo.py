"""GPU parity of hit-and-run, the EI-ZO stages and end-to-end EI-ZO against
the CPU restatement; edge validation, occupancy and k-NN against brute
force."""
import numpy as np
import pytest

from conftest import parse

pytestmark = pytest.mark.gpu


def free_edges(orc, op, lo, hi, count, length, seed, step=0.01):
    """`count` straight edges with collision-free endpoints whose
    checkEdgeCollisionFree discretisation (step 0.01) is free."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    starts, ends = [], []
    while len(starts) < count:
        q = rng.uniform(lo, hi, size=(4096, len(lo))).astype(np.float32)
        free = q[orc.check(op, q.T, orc.RULE_GPU) == 1]
        for q1 in free:
            d = rng.normal(size=len(lo))
            q2 = np.clip(q1 + rng.uniform(*length) * d / np.linalg.norm(d), lo, hi).astype(np.float32)
            if orc.check_edge(op, q1, q2, step) and orc.check(op, q2[:, None], orc.RULE_GPU)[0]:
                starts.append(q1)
                ends.append(q2)
                if len(starts) == count:
                    break
    return np.asfortranarray(np.array(starts).T), np.asfortranarray(np.array(ends).T)


def _box_domain(lo, hi):
    d = len(lo)
    A = np.zeros((2 * d, d), np.float32)
    b = np.zeros(2 * d, np.float32)
    for i in range(d):
        A[2 * i, i], A[2 * i + 1, i] = -1, 1
        b[2 * i], b[2 * i + 1] = -lo[i], hi[i]
    return A, b


def test_hit_and_run_matches_oracle(gpu, orc):
    import torch
    from csdecomp_amd import capi
    A = np.array([[-1, 0], [1, 0], [0, -1], [0, 1], [1, 1]], np.float32)
    b = np.array([0, 6, 0, 6, 11.25], np.float32)
    n = 4096
    dA = torch.from_numpy(A.reshape(-1)).to(gpu)
    db = torch.from_numpy(b).to(gpu)
    nf = torch.tensor([5], dtype=torch.int32, device=gpu)
    st = torch.tensor([1.0, 1.0], device=gpu)
    out = torch.empty(n * 2, device=gpu)
    capi.hit_and_run(dA, db, nf, 5, 2, 1, st, n, 50, 1337, 3, out)
    torch.cuda.synchronize()
    g = out.cpu().numpy().reshape(n, 2)
    ref = orc.hit_and_run(A, b, [1, 1], n, 50, seed=1337, stream=0, round_=3)
    # same Philox stream: only libm-vs-device-math rounding separates them
    assert np.abs(g - ref).max() < 1e-3
    assert np.mean(np.abs(g - ref).max(1) < 1e-4) > 0.99
    assert np.all(g @ A.T <= b + 1e-4)


def test_uniform_sample_api(gpu, csd):
    P1, P2 = csd.HPolyhedron(), csd.HPolyhedron()
    P1.MakeBox(np.array([-1, -1, -1], np.float32), np.array([1, 1, 1], np.float32))
    P2.MakeBox(np.array([0, 0, 0], np.float32), np.array([2, 0.5, 4], np.float32))
    starts = np.array([[0, 0, 0], [1, 0.25, 2]], np.float32).T
    out = csd.UniformSampleInHPolyhedraCuda([P1, P2], starts, 3000, 40)
    assert len(out) == 2 and out[0].shape == (3, 3000)
    for P, s in zip((P1, P2), out):
        assert all(P.PointInSet(s[:, k], 1e-5) for k in range(0, 3000, 7))
    np.testing.assert_allclose(out[1].mean(1), [1, 0.25, 2], atol=0.08)
    assert not np.allclose(out[0][:, 0], out[1][:, 0])
    with pytest.raises(RuntimeError):
        csd.UniformSampleInHPolyhedraCuda([P1], np.array([[5.0, 0, 0]], np.float32).T, 10, 5)


def bisect_outcomes(orc, op, p, lo, steps, band=1e-5):
    """Every result the bisection of cuda_set_builder_utils.cu:239-319 can
    reach when decisions at midpoints within `band` of contact (FP64 signed
    distance) may go either way; float32 arithmetic as on the device."""
    half = np.float32(0.5)
    states = [(p.astype(np.float32), lo.astype(np.float32), p.astype(np.float32))]
    for _ in range(steps):
        nxt = []
        for hi, lo_, sc in states:
            mid = ((hi + lo_) * half).astype(np.float32)
            col = orc.check(op, mid[:, None], orc.RULE_GPU)[0] == 0
            sd = orc.min_sd(op, mid[:, None])[0]
            for c in ({True, False} if abs(sd) < band else {col}):
                nxt.append((mid, lo_, mid) if c else (hi, mid, sc))
        states = nxt
    return [s[2] for s in states]


def test_bisection_stage_parity(gpu, orc):
    """Identical particle sets -> identical optimised collisions, except where
    a midpoint falls in the contact band (then any reachable outcome)."""
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    op = orc.Plant(mp)
    q = scenes.uniform_configs(77, 4000, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    col = q[:, orc.check(op, q, orc.RULE_GPU) == 0][:, :1000]
    ls = np.zeros(7, np.float32)
    le = np.full(7, 0.3, np.float32)
    n = col.shape[1]
    dp = capi.DevicePlant(mp)
    dpart = torch.from_numpy(np.ascontiguousarray(col.T)).to(gpu)
    dls, dle = torch.from_numpy(ls).to(gpu), torch.from_numpy(le).to(gpu)
    seg = torch.zeros(n, dtype=torch.int32, device=gpu)
    scol = torch.empty(n * 7, device=gpu)
    proj = torch.empty(n * 7, device=gpu)
    capi.bisect(dp, dpart, dls, dle, seg, n, 10, scol, proj)
    torch.cuda.synchronize()
    scol, proj = scol.cpu().numpy().reshape(n, 7), proj.cpu().numpy().reshape(n, 7)
    exact = 0
    for k in range(n):
        pr, _ = orc.project(col[:, k], ls, le)
        np.testing.assert_allclose(proj[k], pr, atol=1e-6)
        ref = orc.bisect(op, col[:, k], proj[k], 10)
        if np.array_equal(ref, scol[k]):
            exact += 1
            continue
        outs = bisect_outcomes(orc, op, col[:, k], proj[k], 10)
        assert any(np.array_equal(o, scol[k]) for o in outs), (k, scol[k], ref)
    assert exact >= 0.95 * n


def test_face_stage_parity(gpu, orc):
    import torch
    from csdecomp_amd import capi
    rng = np.random.default_rng(8)
    dim, n, cap = 7, 3000, 64
    ls, le = rng.uniform(-0.2, 0.2, dim).astype(np.float32), rng.uniform(-0.2, 0.2, dim).astype(np.float32)
    proj = (ls + rng.uniform(0, 1, (n, 1)) * (le - ls)).astype(np.float32)
    scol = (proj + rng.normal(0, 0.5, (n, dim))).astype(np.float32)
    A0, b0 = _box_domain(np.full(dim, -3.0), np.full(dim, 3.0))
    A = np.zeros((cap, dim), np.float32)
    b = np.zeros(cap, np.float32)
    A[:14], b[:14] = A0, b0
    added_ref, nf_ref = orc.place_faces(A, b, 14, cap, scol, proj, ls, le, 0.01, 10)
    dA = torch.zeros(cap * dim, device=gpu)
    dA[:14 * dim] = torch.from_numpy(A0.reshape(-1))
    db = torch.zeros(cap, device=gpu)
    db[:14] = torch.from_numpy(b0)
    dnf = torch.tensor([14], dtype=torch.int32, device=gpu)
    added = torch.zeros(1, dtype=torch.int32, device=gpu)
    capi.place_faces(dA, db, dnf, cap, dim, torch.from_numpy(scol).to(gpu), torch.from_numpy(proj).to(gpu),
                     n, torch.from_numpy(ls).to(gpu), torch.from_numpy(le).to(gpu), 0.01, 10, added)
    torch.cuda.synchronize()
    assert int(added.item()) == added_ref == 10 and int(dnf.item()) == nf_ref
    gA = dA.cpu().numpy().reshape(cap, dim)[:nf_ref]
    gb = db.cpu().numpy()[:nf_ref]
    np.testing.assert_allclose(gA, A[:nf_ref], rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(gb, b[:nf_ref], rtol=1e-4, atol=1e-6)


def _boxes():
    from csdecomp_amd import scenes
    p = scenes.boxes_in_corners()
    t = p.get_kinematic_tree()
    dom = None
    import csdecomp_amd as csd
    dom = csd.HPolyhedron()
    dom.MakeBox(t.get_position_lower_limits(), t.get_position_upper_limits())
    return p, p.get_minimal_plant(), dom


@pytest.mark.parametrize("case", [
    # edge_inflation_zero_order_test.cpp FCItest1, FCItest2, FCItest3
    ([0, 0], [1, 0], [[1.8, 0.2]], [[1.8, 0.4]]),
    ([-0.5, 0], [0, 0.5], [[0.0, 0.0]], [[0.0, 0.8]]),
    ([-0.5, 1.75], [0.5, 1.75], [[-1.75, 1.75], [1.75, 1.75]], []),
])
def test_eizo_boxes_in_corners(gpu, csd, orc, case):
    ls, le, inside_pts, outside_pts = case
    p, mp, dom = _boxes()
    opt = csd.EizoOptions()
    opt.num_particles = 1000
    rg = list(p.get_scene_inspector().robot_geometry_ids)
    mpoly, info = csd.EizoCuda(np.array(ls, np.float32), np.array(le, np.float32), None, 0.01, dom,
                               mp, rg, opt)
    P = csd.HPolyhedron(mpoly)
    assert info.converged
    assert P.PointInSet(np.array(ls, np.float32)) and P.PointInSet(np.array(le, np.float32))
    for x in inside_pts:
        assert P.PointInSet(np.array(x, np.float32))
    for x in outside_pts:
        assert not P.PointInSet(np.array(x, np.float32))
    # certified fraction in collision < 1.1 eps over 1000 uniform samples
    c = P.ChebyshevCenter()
    s = csd.UniformSampleInHPolyhedraCuda([P], c[:, None], 1000, 50)[0]
    frac = 1 - np.mean(csd.CheckCollisionFreeCuda(s, mp))
    assert frac < opt.epsilon * 1.1


def test_eizo_voxels(gpu, csd):
    # FCI2dVoxelTest1 (edge_inflation_zero_order_test.cpp:498-586)
    p, mp, dom = _boxes()
    opt = csd.EizoOptions()
    opt.num_particles = 1000
    vox = csd.Voxels(np.array([[-0.55, -1.50, 1.90], [2.05, 1.80, 1.90], [0, 0, 0]], np.float32))
    rg = list(p.get_scene_inspector().robot_geometry_ids)
    inf = csd.CudaEdgeInflator(mp, rg, opt, dom)
    P = inf.inflateEdge(np.array([-0.5, 1.75], np.float32), np.array([0.5, 1.75], np.float32), vox,
                        0.05, False)
    assert P.PointInSet(np.array([-0.5, 1.75], np.float32))
    assert not P.PointInSet(np.array([-1.75, 1.75], np.float32))
    assert P.PointInSet(np.array([1.75, 1.8], np.float32))


def test_eizo_batched_7dof(gpu, csd, orc):
    """Batched inflation of several 7-DOF edges: containment, determinism and
    certification against the exact checker."""
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    op = orc.Plant(mp)
    ls, le = free_edges(orc, op, lo, hi, 6, (0.1, 0.4), seed=21)
    R = ls.shape[1]
    assert R == 6
    dom = csd.HPolyhedron()
    dom.MakeBox(lo, hi)
    opt = csd.EizoOptions()
    opt.epsilon = 0.01
    opt.num_particles = 2000
    opt.bisection_steps = 9
    opt.max_iterations = 60
    rg = list(pl.getRobotGeometryIds())
    inf = csd.CudaEdgeInflator(mp, rg, opt, dom, R, 10)
    regions, infos = inf.inflateEdges(ls, le, None, 0.0, [])
    regions2, _ = inf.inflateEdges(ls, le, None, 0.0, [])
    for r in range(R):
        P = regions[r]
        assert infos[r].converged
        assert P.PointInSet(ls[:, r], 1e-5) and P.PointInSet(le[:, r], 1e-5)
        np.testing.assert_array_equal(P.A(), regions2[r].A())  # deterministic
        c = ((ls[:, r] + le[:, r]) / 2).astype(np.float32)
        s = csd.UniformSampleInHPolyhedraCuda([P], c[:, None], 3000, 50, seed=99 + r)[0]
        frac = 1 - np.mean(csd.CheckCollisionFreeCuda(s, mp))
        assert frac < 2.5 * opt.epsilon, (r, frac)


def test_edges_vs_oracle(gpu, csd, orc):
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    op = orc.Plant(mp)
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    ls, le = free_edges(orc, op, lo, hi, 60, (0.05, 0.5), seed=41, step=0.05)
    q = scenes.uniform_configs(41, 80, lo, hi)
    rng = np.random.default_rng(3)
    nodes = np.concatenate([ls, le, q], axis=1)
    N = nodes.shape[1]
    edges = [(i, 60 + i) for i in range(60)]  # free by construction
    edges += [(i, j) for i in range(N) for j in rng.choice(N, 3, replace=False) if i != j]
    edges = np.array(edges, np.int32)
    dp = capi.DevicePlant(mp)
    dn = torch.from_numpy(np.ascontiguousarray(nodes.T)).to(gpu)
    de = torch.from_numpy(edges).to(gpu)
    out = torch.empty(len(edges), dtype=torch.uint8, device=gpu)
    capi.validate_edges(dp, dn, N, de, len(edges), 0.05, 0, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = np.array([orc.check_edge(op, nodes[:, i], nodes[:, j], 0.05) for i, j in edges], np.uint8)
    assert ref[:60].all() and not ref.all()
    for e in np.nonzero(got != ref)[0]:
        # a mismatch is only allowed if a checked point lies in the contact band
        q1, q2 = nodes[:, edges[e][0]], nodes[:, edges[e][1]]
        n = orc.edge_num_steps(q1, q2, 0.05)
        t = np.array([np.float32(k / (1.0 * n)) for k in range(n)], np.float32)
        pts = (t[None, :] * q2[:, None] + (np.float32(1) - t)[None, :] * q1[:, None]).astype(np.float32)
        assert np.abs(orc.min_sd(op, pts)).min() < 1e-5, e
    assert csd.CheckEdgeCollisionFree(nodes[:, 0], nodes[:, 0], mp)


def _edge_points(q1, q2, step, rule):
    """The points of one edge under each rule, in the float32 arithmetic the
    edge kernel uses (sequential sum of squares, t per rule, q = t q2 +
    (1 - t) q1 rounded per operation)."""
    d = np.float32(0)
    for k in range(len(q1)):
        e = np.float32(q2[k] - q1[k])
        d = np.float32(d + np.float32(e * e))
    dist = np.float32(np.sqrt(d))
    r = np.float32(dist / np.float32(step))
    if rule == 0:
        n = int(np.ceil(r))
        t = [np.float32(k / (1.0 * n)) for k in range(n)]
    elif rule == 2:
        n = max(0, int(np.floor(r)) - 2)
        t = [np.float32((k + 1) / (1.0 * n)) for k in range(n)]
    else:
        n = int(np.uint32(np.float32(r + np.float32(2))))
        ts = np.float32(1.0 / (n - 1))
        t = [np.float32(0), np.float32(1)] + [np.float32(ts * np.float32(k - 1)) for k in range(2, n)]
    t = np.asarray(t, np.float32)[:, None]
    return (t * q2[None, :] + (np.float32(1) - t) * q1[None, :]).astype(np.float32)


@pytest.mark.parametrize("rule", [0, 1, 2])
def test_edges_match_their_point_checks(gpu, rule):
    """Every rule: an edge's flag is the AND of the checker's flags at its
    points, bit for bit (the same device arithmetic), over 5,013 edges (not
    a multiple of 64) between free and colliding nodes, with zero-length
    edges, invalid node indices and an absurdly long edge."""
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    dp = capi.DevicePlant(mp)
    cand = scenes.uniform_configs(77, 20000, lo, hi).T.copy()  # [n][7]
    dc = torch.from_numpy(cand).to(gpu)
    fl = torch.empty(len(cand), dtype=torch.uint8, device=gpu)
    capi.check_collision_free(dp, dc, len(cand), 7, 1, fl)
    torch.cuda.synchronize()
    fl = fl.cpu().numpy()
    free_nodes, col_nodes = cand[fl == 1][:300], cand[fl == 0][:300]
    assert len(free_nodes) >= 100
    nodes = np.concatenate([free_nodes, col_nodes]).astype(np.float32)
    nodes = np.concatenate([nodes, (nodes[:1] + np.float32(1e7))])  # far away: too many points
    N = len(nodes)
    rng = np.random.default_rng(rule)
    nf = len(free_nodes)
    E = 5013
    a = np.where(rng.random(E) < 0.7, rng.integers(0, nf, E), rng.integers(0, N - 1, E))
    # short edges between free nodes: nearest other free node
    b = rng.integers(0, N - 1, E)
    near = rng.random(E) < 0.5
    fa = a < nf
    for i in np.nonzero(near & fa)[0]:
        d2 = ((free_nodes - nodes[a[i]]) ** 2).sum(1)
        d2[a[i]] = np.inf
        b[i] = int(np.argmin(d2))
    edges = np.stack([a, b], 1).astype(np.int32)
    edges[10] = (3, 3)  # zero length
    edges[11] = (-1, 4)  # invalid
    edges[12] = (5, N)  # invalid
    edges[13] = (0, N - 1)  # too long to discretise
    step = 0.05
    de = torch.from_numpy(edges).to(gpu)
    dn = torch.from_numpy(nodes).to(gpu)
    out = torch.full((E,), 7, dtype=torch.uint8, device=gpu)
    capi.validate_edges(dp, dn, N, de, E, step, rule, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert got[11] == 0 and got[12] == 0 and got[13] == 0
    pts, owner = [], []
    for e in range(E):
        if e in (11, 12, 13):
            continue
        p = _edge_points(nodes[edges[e, 0]], nodes[edges[e, 1]], step, rule)
        pts.append(p)
        owner.append(np.full(len(p), e))
    pts = np.concatenate(pts)
    owner = np.concatenate(owner)
    dq = torch.from_numpy(np.ascontiguousarray(pts)).to(gpu)
    pf = torch.empty(len(pts), dtype=torch.uint8, device=gpu)
    capi.check_collision_free(dp, dq, len(pts), 7, 1, pf)
    torch.cuda.synchronize()
    pf = pf.cpu().numpy()
    ref = np.ones(E, np.uint8)
    np.minimum.at(ref, owner, pf)
    ref[[11, 12, 13]] = 0
    np.testing.assert_array_equal(got, ref)
    assert 0.1 < ref.mean() < 0.9, ref.mean()  # both verdicts well represented


def test_visibility_graph_reference_pattern(gpu, csd):
    # cuda_visibility_graph_test.cpp:132-155
    p, mp, _ = _boxes()
    nodes = np.array([[0, 1.5, 0, -1.5, 0], [0, 0, 1.5, 0, -1.5]], np.float32)
    A = csd.VisibilityGraph(nodes, mp, 0.1, 1000000)
    exp = {(0, 1): 1, (0, 2): 1, (0, 3): 1, (0, 4): 1, (1, 2): 0, (1, 3): 1, (1, 4): 0, (2, 3): 0,
           (2, 4): 1, (3, 4): 0}
    for (i, j), v in exp.items():
        assert A[i, j] == v, (i, j)
    B = csd.VisibilityGraph(nodes, mp, 0.1, 100)
    assert (A != B).nnz == 0


def test_occupancy_vs_brute_force(gpu, orc):
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1000, 16, 0)
    mp = pl.getMinimalPlant()
    op = orc.Plant(mp)
    rg = list(pl.getRobotGeometryIds())
    q = scenes.uniform_configs(51, 20, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    res, n = 1.0 / 16, 16  # 16^3 grid over [-1, 1]^3
    dp = capi.DevicePlant(mp)
    grid = torch.zeros(n ** 3, dtype=torch.uint8, device=gpu)
    capi.occupancy(dp, torch.from_numpy(np.ascontiguousarray(q)).to(gpu), q.shape[1], 1, q.shape[1], rg,
                   [0, 0, 0], [2, 2, 2], res, grid)
    torch.cuda.synchronize()
    got = grid.cpu().numpy()
    idx = np.arange(n)
    cen = (np.float64(0) - 2 / 2.0 + (idx * 2 * np.float32(res)).astype(np.float32) + res).astype(np.float32)
    vx = np.array([[cen[ix], cen[iy], cen[iz]] for iz in idx for iy in idx for ix in idx], np.float32)
    vr = np.float32(np.sqrt(3.0) * res)
    flags = orc.voxel_flags(op, q, vx.T, vr, rg, orc.RULE_GPU)  # [N][V], 1 = free
    ref = (flags == 0).any(0).astype(np.uint8)
    assert ref.sum() > 50
    # a voxel may differ only where some configuration's robot is within
    # 1e-5 m of its sphere: a radius 1e-5 m smaller and one 1e-5 m larger
    # then give the oracle different occupancies
    lo = (orc.voxel_flags(op, q, vx.T, vr - np.float32(1e-5), rg, orc.RULE_GPU) == 0).any(0)
    hi = (orc.voxel_flags(op, q, vx.T, vr + np.float32(1e-5), rg, orc.RULE_GPU) == 0).any(0)
    diff = got != ref
    assert not np.any(diff & (lo == hi)), np.argwhere(diff & (lo == hi))[:5]
    assert diff.sum() <= 4, np.argwhere(diff)[:5]


def test_knn_vs_brute_force(gpu):
    import torch
    from csdecomp_amd import capi
    rng = np.random.default_rng(2)
    nodes = rng.uniform(-np.pi, np.pi, (3000, 7)).astype(np.float32)
    nodes[17] = nodes[5]  # exact duplicate is skipped like the node itself
    k = 10
    dn = torch.from_numpy(nodes).to(gpu)
    idx = torch.empty(3000 * k, dtype=torch.int32, device=gpu)
    d2 = torch.empty(3000 * k, device=gpu)
    capi.knn(dn, 3000, 7, k, idx, d2)
    torch.cuda.synchronize()
    idx = idx.cpu().numpy().reshape(3000, k)
    for i in range(0, 3000, 37):
        dd = ((nodes - nodes[i]) ** 2).sum(1)
        dd[dd <= np.finfo(np.float32).eps] = np.inf
        order = np.lexsort((np.arange(3000), dd))[:k]
        assert list(order) == list(idx[i]), i
    assert 17 not in idx[5] and 5 not in idx[17]


def _train_surrogate(gpu, pl, ls=None, le=None, steps=300, n=262144, seed=5):
    from csdecomp_amd import surrogate
    return surrogate.train_surrogate(pl, ls, le, n=n, steps=steps, seed=seed, device=gpu)[0]


def test_guided_eizo_certifies(gpu, csd, orc):
    """Pilot rounds and surrogate-ranked particles: every certified region
    contains its segment and passes an independent collision-fraction check;
    runs are reproducible."""
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    op = orc.Plant(mp)
    ls, le = free_edges(orc, op, lo, hi, 6, (0.1, 0.4), seed=23)
    R = ls.shape[1]
    dom = csd.HPolyhedron()
    dom.MakeBox(lo, hi)
    opt = csd.EizoOptions()
    opt.epsilon = 0.01
    opt.num_particles = 2000
    opt.bisection_steps = 9
    opt.max_iterations = 80
    rg = list(pl.getRobotGeometryIds())
    model = _train_surrogate(gpu, pl, ls, le)
    results = {}
    for mode in ("plain", "pilot", "guided", "guided_wide"):
        inf = csd.CudaEdgeInflator(mp, rg, opt, dom, R, 10)
        if mode == "pilot":
            inf.set_guidance(pilot_samples=2048, pilot_factor=10.0)
        elif mode.startswith("guided"):
            # 32 of the ranked particles, or all 1,024 ranked (the bitonic order)
            inf.set_guidance(pilot_samples=2048, pilot_factor=10.0, surrogate=model,
                             keep_particles=32 if mode == "guided" else 1024, line_points=4)
        regions, infos = inf.inflateEdges(ls, le, None, 0.0, [])
        again, _ = inf.inflateEdges(ls, le, None, 0.0, [])
        results[mode] = infos
        for r in range(R):
            P = regions[r]
            assert P.PointInSet(ls[:, r], 1e-5) and P.PointInSet(le[:, r], 1e-5)
            np.testing.assert_array_equal(P.A(), again[r].A())
            if infos[r].converged:
                c = ((ls[:, r] + le[:, r]) / 2).astype(np.float32)
                s = csd.UniformSampleInHPolyhedraCuda([P], c[:, None], 3000, 50, seed=99 + r)[0]
                frac = 1 - np.mean(csd.CheckCollisionFreeCuda(s, mp))
                assert frac < 2.5 * opt.epsilon, (mode, r, frac)
        assert sum(i.converged for i in infos) >= R - 1, mode


def test_guidance_argument_checks(gpu, csd):
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    dom = csd.HPolyhedron()
    dom.MakeBox(pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    inf = csd.CudaEdgeInflator(pl.getMinimalPlant(), list(pl.getRobotGeometryIds()), csd.EizoOptions(), dom)
    with pytest.raises((ValueError, RuntimeError)):
        inf.set_guidance(pilot_samples=1000, pilot_factor=0.0)
    with pytest.raises((ValueError, RuntimeError)):
        inf.set_guidance(surrogate=1, keep_particles=0)
    with pytest.raises((ValueError, RuntimeError)):
        inf.set_guidance(surrogate=1, keep_particles=1025)


@pytest.mark.parametrize("scene", ["arm7", "dual14"])
def test_graph_rounds_match_host_rounds(gpu, csd, scene):
    """Device-resident rounds (one graph-captured round per launch, region
    list built on the GPU) give the host-driven rounds' polytopes, round
    counts and sample counts bit for bit: uneven particle budgets, several
    status-read intervals, replays of the captured graph with host-driven
    runs in between, and a second call with fewer regions."""
    from csdecomp_amd import scenes
    if scene == "arm7":
        pl, R = scenes.arm_scene(1001, 48, 16), 6
    else:
        pl, R = scenes.dual_arm_scene(1002, 72, 24), 4
    mp = pl.getMinimalPlant()
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    ls, le = scenes.free_segments(pl, R, 3100, (0.1, 0.6))
    dom = csd.HPolyhedron()
    dom.MakeBox(lo, hi)
    opt = csd.EizoOptions()
    opt.epsilon = 0.01
    opt.num_particles = 2000
    opt.bisection_steps = 9
    opt.max_iterations = 30
    opt.verbose = False
    budgets = [500, 2000, 1000, 1800, 800, 1500][:R]  # <= num_particles (the capacity)
    inf = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), opt, dom, R, 10)

    def run(sync, n):
        inf.set_graph(sync)
        regions, infos = inf.inflateEdges(np.asfortranarray(ls[:, :n]), np.asfortranarray(le[:, :n]),
                                          None, 0.0, budgets[:n])
        return regions, infos

    runs = [run(0, R), run(1, R), run(3, R), run(0, R), run(1, R), run(2, R - 2), run(0, R - 2)]
    for k, (regions, infos) in enumerate(runs):
        ref_regions, ref_infos = runs[0] if len(regions) == R else runs[-1]
        for r in range(len(regions)):
            np.testing.assert_array_equal(regions[r].A(), ref_regions[r].A(), err_msg=f"run {k} region {r}")
            np.testing.assert_array_equal(regions[r].b(), ref_regions[r].b(), err_msg=f"run {k} region {r}")
            got, ref = infos[r], ref_infos[r]
            assert (got.iterations, got.converged, got.samples_drawn, got.faces_added,
                    got.face_capacity_reached) == \
                (ref.iterations, ref.converged, ref.samples_drawn, ref.faces_added,
                 ref.face_capacity_reached), (k, r)
    assert sum(i.converged for i in runs[0][1]) >= 1
    assert max(i.iterations for i in runs[0][1]) >= 2  # several rounds ran


@pytest.mark.parametrize("scene", ["arm7", "dual14"])
def test_sampler_tiles_give_the_same_regions(gpu, csd, scene):
    """The hit-and-run sampler's tiles draw the same samples, so the regions
    match bit for bit: 256-thread workgroups (the default), 1024-thread
    workgroups (set_sampler_tile(0)), the tuned choice, and graph-captured
    rounds, which stage LDS for the whole face capacity: above 64 KB only
    the 1024-thread tile fits (2,000 faces at 7 dims, 1,100 at 14), above
    160 KB the polytope is read from global memory (4,600 and 2,500)."""
    from csdecomp_amd import scenes
    if scene == "arm7":
        pl, R, caps = scenes.arm_scene(1001, 48, 16), 4, (200, 2000, 4600)
    else:
        pl, R, caps = scenes.dual_arm_scene(1002, 72, 24), 3, (200, 1100, 2500)
    mp = pl.getMinimalPlant()
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    ls, le = scenes.free_segments(pl, R, 3200, (0.1, 0.6))
    dom = csd.HPolyhedron()
    dom.MakeBox(lo, hi)
    opt = csd.EizoOptions()
    opt.epsilon = 0.01
    opt.num_particles = 1000
    opt.bisection_steps = 9
    opt.max_iterations = 8
    opt.verbose = False
    budgets = [400, 1000, 700, 900][:R]
    ids = list(pl.getRobotGeometryIds())

    def run(inf, graph):
        inf.set_graph(graph)
        regions, infos = inf.inflateEdges(ls, le, None, 0.0, budgets)
        return [(r.A(), r.b()) for r in regions], [(i.iterations, i.samples_drawn) for i in infos]

    def same(a, b, what):
        for r, ((A1, b1), (A2, b2)) in enumerate(zip(a[0], b[0])):
            np.testing.assert_array_equal(A1, A2, err_msg=f"{what} region {r}")
            np.testing.assert_array_equal(b1, b2, err_msg=f"{what} region {r}")
        assert a[1] == b[1], what

    for cap in caps:
        inf = csd.CudaEdgeInflator(mp, ids, opt, dom, R, 10, cap)
        ref = run(inf, 0)  # host rounds: the tile of each round's largest polytope
        inf.set_sampler_tile(0)
        same(run(inf, 0), ref, f"cap {cap}: 1024-thread tile")
        same(run(inf, 1), ref, f"cap {cap}: graph rounds, 1024-thread tile where it fits")
        inf.set_sampler_tile(2 ** 31 - 1)
        same(run(inf, 1), ref, f"cap {cap}: graph rounds, default tile")
        if cap == caps[0]:
            wide_from = inf.tune_sampler()
            assert wide_from >= 0
            same(run(inf, 0), ref, f"tuned (1024-thread tile from {wide_from} faces)")
        assert max(i for i, _ in ref[1]) >= 2  # several rounds ran


@pytest.mark.parametrize("mode", ["pilot", "ranked", "pruned", "ranked_no_pilot", "dual_pruned"])
def test_guided_graph_rounds_match_host_rounds(gpu, csd, mode):
    """Guided rounds device-resident: the pilot's decision, the pruned
    pilot's predictions and the surrogate ranking run over slot lists built
    on the GPU inside the captured round, and give the host-driven guided
    rounds' polytopes, round counts and sample counts bit for bit, for
    several status-read intervals and a second call with fewer regions."""
    from csdecomp_amd import scenes
    if mode.startswith("dual"):
        pl, R = scenes.dual_arm_scene(1002, 72, 24), 4
    else:
        pl, R = scenes.arm_scene(1001, 48, 16), 6
    mp = pl.getMinimalPlant()
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    ls, le = scenes.free_segments(pl, R, 3100, (0.1, 0.6))
    dom = csd.HPolyhedron()
    dom.MakeBox(lo, hi)
    opt = csd.EizoOptions()
    opt.epsilon = 0.01
    opt.num_particles = 2000
    opt.bisection_steps = 9
    opt.max_iterations = 30
    opt.verbose = False
    opt.track_iteration_information = True
    budgets = [500, 2000, 1000, 1800, 800, 1500][:R]
    model = None if mode == "pilot" else _train_surrogate(gpu, pl, ls, le, steps=300)
    inf = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), opt, dom, R, 10)
    pilot = 0 if mode == "ranked_no_pilot" else 2048
    inf.set_guidance(pilot_samples=pilot, pilot_factor=10.0, surrogate=model, keep_particles=64,
                     line_points=4, prune_below=0.1 if mode.endswith("pruned") else None)

    def run(sync, n):
        inf.set_graph(sync)
        return inf.inflateEdges(np.asfortranarray(ls[:, :n]), np.asfortranarray(le[:, :n]), None,
                                0.0, budgets[:n])

    runs = [run(0, R), run(1, R), run(4, R), run(0, R - 2), run(2, R - 2)]
    for k, (regions, infos) in enumerate(runs):
        ref_regions, ref_infos = runs[0] if len(regions) == R else runs[3]
        for r in range(len(regions)):
            np.testing.assert_array_equal(regions[r].A(), ref_regions[r].A(), err_msg=f"run {k} region {r}")
            np.testing.assert_array_equal(regions[r].b(), ref_regions[r].b(), err_msg=f"run {k} region {r}")
            got, ref = infos[r], ref_infos[r]
            assert (got.iterations, got.converged, got.samples_drawn, got.faces_added) == \
                (ref.iterations, ref.converged, ref.samples_drawn, ref.faces_added), (k, r)
            gi = [(x.num_samples_drawn, x.num_pilot_samples, x.num_pilot_checked,
                   x.num_pilot_collisions, x.num_collisions_updated, x.num_faces_placed)
                  for x in got.iteration_info]
            ri = [(x.num_samples_drawn, x.num_pilot_samples, x.num_pilot_checked,
                   x.num_pilot_collisions, x.num_collisions_updated, x.num_faces_placed)
                  for x in ref.iteration_info]
            assert gi == ri, (k, r)
    infos = runs[0][1]
    assert max(i.iterations for i in infos) >= 2
    rounds = [x for i in infos for x in i.iteration_info]
    if pilot:
        # both pilot outcomes occur: some rounds skip the test, some run it
        assert any(x.num_samples_drawn == 0 for x in rounds)
        assert any(x.num_samples_drawn > 0 for x in rounds)
    if mode.endswith("pruned"):
        assert sum(x.num_pilot_checked for x in rounds) < sum(x.num_pilot_samples for x in rounds)


def test_graph_rounds_argument_checks(gpu, csd):
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    dom = csd.HPolyhedron()
    dom.MakeBox(pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    inf = csd.CudaEdgeInflator(pl.getMinimalPlant(), list(pl.getRobotGeometryIds()), csd.EizoOptions(), dom)
    with pytest.raises((ValueError, RuntimeError)):
        inf.set_graph(-1)
    with pytest.raises((ValueError, RuntimeError)):
        inf.set_graph(65)


def test_guided_plain_fallback(gpu, csd):
    """plain_fallback: regions the guided rounds certify keep the guided
    polytope; the others are exactly what plain EI-ZO gives for that subset
    of edges (same order, same budgets), and their info counts both runs."""
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    lo, hi = pl.getPositionLowerLimits(), pl.getPositionUpperLimits()
    ls, le = scenes.free_segments(pl, 8, 3102, (0.4, 1.0))
    R = ls.shape[1]
    dom = csd.HPolyhedron()
    dom.MakeBox(lo, hi)
    opt = csd.EizoOptions()
    opt.epsilon = 0.01
    opt.num_particles = 1500
    opt.bisection_steps = 9
    opt.max_iterations = 3
    opt.verbose = False
    rg = list(pl.getRobotGeometryIds())
    budgets = [600, 1500, 900, 1200, 700, 1500, 1000, 800]

    def inflator(fallback=None):
        inf = csd.CudaEdgeInflator(mp, rg, opt, dom, R, 10)
        if fallback is not None:
            inf.set_guidance(pilot_samples=2048, pilot_factor=10.0, plain_fallback=fallback)
        return inf

    guided, ginfo = inflator(False).inflateEdges(ls, le, None, 0.0, budgets)
    both, binfo = inflator(True).inflateEdges(ls, le, None, 0.0, budgets)
    redo = [r for r in range(R) if not ginfo[r].converged]
    assert redo, "every edge certified in the guided rounds; the fallback went unused"
    sub = np.asfortranarray(ls[:, redo]), np.asfortranarray(le[:, redo])
    plain, pinfo = inflator().inflateEdges(*sub, None, 0.0, [budgets[r] for r in redo])
    for r in range(R):
        if r in redo:
            k = redo.index(r)
            ref, refinfo = plain[k], pinfo[k]
            assert binfo[r].plain_fallback
            assert binfo[r].converged == refinfo.converged
            assert binfo[r].iterations == ginfo[r].iterations + refinfo.iterations
            assert binfo[r].samples_drawn == ginfo[r].samples_drawn + refinfo.samples_drawn
        else:
            ref = guided[r]
            assert not binfo[r].plain_fallback and binfo[r].converged
        np.testing.assert_array_equal(both[r].A(), ref.A())
        np.testing.assert_array_equal(both[r].b(), ref.b())
        assert both[r].PointInSet(ls[:, r], 1e-5) and both[r].PointInSet(le[:, r], 1e-5)
