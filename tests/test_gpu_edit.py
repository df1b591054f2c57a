"""GPU parity of region editing (csd_edit_regions, EditRegionsCuda) against
the oracle: projections, bisected collisions (exact outside the contact
band) and the faces each region gains, plus the C-ABI's status codes."""
import numpy as np
import pytest

from test_gpu_eizo import free_edges
from test_oracle_edit import COL, LE, LS, VOX, _box, _in_set, reference_case

pytestmark = pytest.mark.gpu


def _dom(csd, plant):
    d = csd.HPolyhedron()
    d.MakeBox(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    return d


def test_reference_scenario_gpu(gpu, orc, csd):
    """planning/tests/cuda_edit_regions_test.cpp:74-216 through the Python
    names, checked against its assertions and against the oracle."""
    plant, collisions, seg = reference_case()
    mp = plant.getMinimalPlant()
    rg = plant.getSceneInspector().robot_geometry_ids
    dom = _dom(csd, plant)
    opts = csd.EditRegionsOptions()
    assert (opts.configuration_margin, opts.bisection_steps, opts.verbose) == (
        pytest.approx(0.01), 9, False)
    regions, (proj, opt) = csd.EditRegionsCuda(collisions, seg, LS, LE, [dom] * 3, mp, rg, VOX,
                                               0.05, opts)
    assert len(regions) == 3 and proj.shape == (2, 12) and opt.shape == (2, 12)
    for r, reg in enumerate(regions):
        assert reg.PointInSet(LS[:, r]) and reg.PointInSet(LE[:, r])
        for c in range(12):
            assert not reg.PointInSet(opt[:, c])
        for c in range(4):
            assert not reg.PointInSet(COL[:, c])
    op = orc.Plant(mp)
    A, b = _box(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    ref, rproj, ropt = orc.edit_regions(op, collisions, seg, LS, LE, [(A, b)] * 3, rg, VOX, 0.05)
    np.testing.assert_array_equal(proj, rproj)
    np.testing.assert_array_equal(opt, ropt)
    for reg, (Ar, br) in zip(regions, ref):
        assert reg.A().shape == Ar.shape
        np.testing.assert_allclose(reg.A(), Ar, atol=1e-6)
        np.testing.assert_allclose(reg.b(), br, atol=1e-6)
    # a single free midpoint leaves the collision where it was
    opts.bisection_steps = 1
    _, (_, opt1) = csd.EditRegionsCuda(collisions, seg, LS, LE, [dom] * 3, mp, rg, VOX, 0.05,
                                       opts)
    assert np.linalg.norm(opt1[:, 1] - COL[:, 1]) <= 1e-6


def _bisect_outcomes(orc, op, p, lo, steps, vox, radius, rg, band=1e-5):
    """Results reachable when checks within `band` of contact go either way."""
    half = np.float32(0.5)
    states = [(p.astype(np.float32), lo.astype(np.float32), p.astype(np.float32))]
    for _ in range(steps):
        nxt = []
        for hi, lo_, sc in states:
            mid = ((hi + lo_) * half).astype(np.float32)
            col = orc.check_voxels(op, mid[:, None], vox, radius, rg)[0] == 0
            sd = orc.min_sd(op, mid[:, None], vox, radius, rg)[0]
            for c in ({True, False} if abs(sd) < band else {col}):
                nxt.append((mid, lo_, mid) if c else (hi, mid, sc))
        states = nxt
    return [s[2] for s in states]


def _arm_case(orc, R=24, per=40, seed=5):
    """7-DOF arm, 48 boxes + 16 spheres, 32 voxel spheres; region r is the box
    around free segment r widened by 0.4 rad; up to `per` colliding samples
    per region (none for every fifth), shuffled."""
    from csdecomp_amd import scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    op = orc.Plant(mp)
    lo = np.float32(pl.getPositionLowerLimits())
    hi = np.float32(pl.getPositionUpperLimits())
    ls, le = free_edges(orc, op, lo, hi, R, (0.2, 0.8), seed)
    rng = np.random.default_rng(seed)
    vox = rng.uniform(-0.8, 0.8, (3, 32)).astype(np.float32)
    rg = list(pl.getSceneInspector().robot_geometry_ids)
    regions, cols, seg = [], [], []
    for r in range(R):
        blo = np.maximum(lo, np.minimum(ls[:, r], le[:, r]) - 0.4)
        bhi = np.minimum(hi, np.maximum(ls[:, r], le[:, r]) + 0.4)
        regions.append(_box(blo, bhi))
        if r % 5 == 4:
            continue
        q = rng.uniform(blo, bhi, (2048, 7)).astype(np.float32)
        c = q[orc.check_voxels(op, q.T, vox, 0.02, rg) == 0][:per]
        cols.append(c)
        seg += [r] * len(c)
    cols = np.concatenate(cols)
    perm = rng.permutation(len(seg))
    return pl, mp, op, ls, le, vox, rg, regions, cols[perm], np.asarray(seg, np.uint32)[perm]


def _run(gpu, mp, cols, seg, ls, le, regions, vox, radius, rg, margin, steps, cap=None,
         offsets=None):
    """csd_edit_regions through the C-ABI. The collisions are grouped by
    region (stable) first; returns A, b, nf, status, scol, proj and the
    grouped collisions and segment indices."""
    import torch
    from csdecomp_amd import capi
    R, dim = len(regions), ls.shape[0]
    order = np.argsort(seg, kind="stable")
    cols, seg = np.asarray(cols, np.float32).reshape(-1, dim)[order], np.asarray(seg, np.uint32)[order]
    counts = np.bincount(seg, minlength=R) if len(seg) else np.zeros(R, np.int64)
    if offsets is None:
        offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    if cap is None:
        cap = max(A.shape[0] + int(n) for (A, _), n in zip(regions, counts))
    A = np.zeros((R, cap, dim), np.float32)
    b = np.zeros((R, cap), np.float32)
    nf = np.zeros(R, np.int32)
    for r, (Ar, br) in enumerate(regions):
        A[r, :len(br)], b[r, :len(br)], nf[r] = Ar, br, len(br)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(gpu)  # noqa: E731
    C = len(seg)
    dA, db, dnf = t(A), t(b), t(nf)
    dsc = torch.empty(max(C, 1) * dim, device=gpu)
    dpr = torch.empty(max(C, 1) * dim, device=gpu)
    dst = torch.full((R,), -99, dtype=torch.int32, device=gpu)
    dvox = t(vox.T) if vox is not None else None
    dplant = capi.DevicePlant(mp)
    capi.edit_regions(dplant, t(cols) if C else None,
                      t(seg.view(np.int32)) if C else None, t(offsets), C, t(ls.T), t(le.T), R,
                      dA, db, dnf, cap, margin, steps, dsc, dpr, dst, voxels=dvox,
                      num_voxels=0 if vox is None else vox.shape[1], radius=radius,
                      robot_geoms=rg)
    torch.cuda.synchronize()
    return (dA.cpu().numpy(), db.cpu().numpy(), dnf.cpu().numpy(), dst.cpu().numpy(),
            dsc.cpu().numpy()[:C * dim].reshape(C, dim),
            dpr.cpu().numpy()[:C * dim].reshape(C, dim), cols, seg)


def test_edit_regions_parity_arm(gpu, orc):
    pl, mp, op, ls, le, vox, rg, regions, cols, seg = _arm_case(orc)
    margin, steps = 0.01, 9
    A, b, nf, st, scol, proj, cols, seg = _run(gpu, mp, cols, seg, ls, le, regions, vox, 0.02,
                                               rg, margin, steps)
    exact = 0
    for k in range(len(seg)):
        s = seg[k]
        pr, _ = orc.project(cols[k], ls[:, s], le[:, s])
        np.testing.assert_array_equal(proj[k], pr)  # same float operations, same bits
        ref = orc.bisect_voxels(op, cols[k], proj[k], steps, vox, 0.02, rg)
        if np.array_equal(ref, scol[k]):
            exact += 1
            continue
        outs = _bisect_outcomes(orc, op, cols[k], proj[k], steps, vox, 0.02, rg)
        assert any(np.array_equal(o, scol[k]) for o in outs), (k, scol[k], ref)
    assert exact >= 0.95 * len(seg)
    placed = 0
    for r, (Ar, br) in enumerate(regions):
        m = seg == r
        A2, b2, added = orc.edit_region(Ar, br, scol[m], proj[m], ls[:, r], le[:, r], margin)
        assert st[r] == added and nf[r] == A2.shape[0], (r, st[r], added)
        np.testing.assert_allclose(A[r, :nf[r]], A2, atol=1e-6)
        np.testing.assert_allclose(b[r, :nf[r]], b2, atol=1e-6)
        if not m.any():
            assert added == 0
        # the segment stays inside, every bisected collision is cut off
        assert _in_set(A2, b2, ls[:, r]) and _in_set(A2, b2, le[:, r])
        placed += added
    assert placed > len(regions)


def test_edit_regions_status_codes(gpu, orc):
    plant, collisions, seg = reference_case()
    mp = plant.getMinimalPlant()
    rg = plant.getSceneInspector().robot_geometry_ids
    box = _box(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    seg = np.asarray(seg, np.uint32)
    full = _run(gpu, mp, collisions.T, seg, LS, LE, [box] * 3, VOX, 0.05, rg, 0.01, 9)
    assert (full[3] > 1).all()
    # room for one face: status -2, the first face kept and equal to the full run's
    A, b, nf, st, *_ = _run(gpu, mp, collisions.T, seg, LS, LE, [box] * 3, VOX, 0.05, rg,
                              0.01, 9, cap=5)
    assert (st == -2).all() and (nf == 5).all()
    np.testing.assert_array_equal(A[:, :5], full[0][:, :5])
    np.testing.assert_array_equal(b[:, :5], full[1][:, :5])
    # a region without collisions and an empty batch are left untouched
    A, b, nf, st, *_ = _run(gpu, mp, collisions.T[:4], seg[:4], LS, LE, [box] * 3, VOX, 0.05,
                              rg, 0.01, 9)
    assert st[0] > 0 and st[1] == 0 and st[2] == 0 and nf[1] == 4 and nf[2] == 4
    np.testing.assert_array_equal(A[1, :4], box[0])
    A, b, nf, st, *_ = _run(gpu, mp, np.zeros((0, 2), np.float32), np.zeros(0, np.uint32), LS,
                              LE, [box] * 3, None, 0.0, rg, 0.01, 9)
    assert (st == 0).all() and (nf == 4).all()
    # more than 4096 collisions for one region: status -1, region untouched
    many = np.tile(COL.T[:1], (4097, 1))
    A, b, nf, st, *_ = _run(gpu, mp, many, np.zeros(4097, np.uint32), LS, LE, [box] * 3, VOX,
                              0.05, rg, 0.01, 2, cap=8)
    assert st[0] == -1 and nf[0] == 4 and st[1] == 0
    # offsets that decrease or run past the batch: status -4, regions untouched
    A, b, nf, st, *_ = _run(gpu, mp, collisions.T, seg, LS, LE, [box] * 3, VOX, 0.05, rg, 0.01,
                            9, offsets=np.array([0, 4, 2, 13], np.int32))
    assert st[0] > 0 and st[1] == -4 and st[2] == -4 and nf[1] == 4 and nf[2] == 4


def test_bisect_split_levels_agree(gpu, orc):
    """csd_bisect spreads small batches over 4 or 2 waves per 64 particles and
    large ones over 1: the first 1000 particles bisect to the same bits in
    batches of 1000 (4 waves), 140,000 (2) and 300,000 (1)."""
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1001, 48, 16)
    dp = capi.DevicePlant(pl.getMinimalPlant())
    q = scenes.uniform_configs(91, 300_000, pl.getPositionLowerLimits(),
                               pl.getPositionUpperLimits())
    part = torch.from_numpy(np.ascontiguousarray(q.T)).to(gpu)
    ls = torch.zeros(7, device=gpu)
    le = torch.full((7,), 0.5, device=gpu)
    seg = torch.zeros(300_000, dtype=torch.int32, device=gpu)
    res = []
    for n in (1000, 140_000, 300_000):
        scol = torch.empty(n * 7, device=gpu)
        proj = torch.empty(n * 7, device=gpu)
        capi.bisect(dp, part, ls, le, seg, n, 9, scol, proj)
        torch.cuda.synchronize()
        res.append((scol[:7000].cpu().numpy(), proj[:7000].cpu().numpy()))
    for sc, pr in res[1:]:
        np.testing.assert_array_equal(sc, res[0][0])
        np.testing.assert_array_equal(pr, res[0][1])
    assert not np.array_equal(res[0][0], q.T[:1000].reshape(-1))  # some points moved


def test_edit_regions_many_regions(gpu, orc):
    """A larger batch: the dual-arm scene's 14 dimensions, 96 regions."""
    from csdecomp_amd import scenes
    pl = scenes.dual_arm_scene(7)
    mp = pl.getMinimalPlant()
    op = orc.Plant(mp)
    lo = np.float32(pl.getPositionLowerLimits())
    hi = np.float32(pl.getPositionUpperLimits())
    R = 96
    ls, le = free_edges(orc, op, lo, hi, R, (0.1, 0.4), 11)
    rng = np.random.default_rng(12)
    rg = list(pl.getSceneInspector().robot_geometry_ids)
    regions, cols, seg = [], [], []
    for r in range(R):
        blo = np.maximum(lo, np.minimum(ls[:, r], le[:, r]) - 0.3)
        bhi = np.minimum(hi, np.maximum(ls[:, r], le[:, r]) + 0.3)
        regions.append(_box(blo, bhi))
        q = rng.uniform(blo, bhi, (1024, 14)).astype(np.float32)
        c = q[orc.check(op, q.T, orc.RULE_GPU) == 0][:24]
        cols.append(c)
        seg += [r] * len(c)
    cols, seg = np.concatenate(cols), np.asarray(seg, np.uint32)
    A, b, nf, st, scol, proj, cols, seg = _run(gpu, mp, cols, seg, ls, le, regions, None, 0.0,
                                               rg, 0.02, 9)
    assert (st >= 0).all()
    for r, (Ar, br) in enumerate(regions):
        m = seg == r
        A2, b2, added = orc.edit_region(Ar, br, scol[m], proj[m], ls[:, r], le[:, r], 0.02)
        assert st[r] == added and nf[r] == A2.shape[0]
        np.testing.assert_allclose(A[r, :nf[r]], A2, atol=1e-6)
        np.testing.assert_allclose(b[r, :nf[r]], b2, atol=1e-6)
