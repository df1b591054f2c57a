"""EI-ZO round parity: the GPU's batched rounds against the oracle's
restatement of one round (cuda_edge_inflation_zero_order.cu:109-306), stage
by stage on the GPU's own data.

The GPU's sampler uses the hardware log, sin/cos and reciprocal, so its
samples agree with the oracle's libm chains only to ~1e-4 (checked below).
To compare the rest of the round exactly, the round's samples are
regenerated bit for bit with csd_hit_and_run_segments (the round's sampler
on its own) and the oracle replays the round on them:
  1. collision flags of every sample (FP32 mirror, band-aware),
  2. the unadaptive test: collisions among the first M_k (a13),
  3. which colliding samples become particles: the first ones in sample
     order, up to the budget (a14),
  4. projection and bisection of each particle (a15, a16),
  5. FP64 face placement on the GPU's bisected points (a17), to 1e-4
     relative.
Workloads: the 7-DOF arm with 48 boxes + 16 spheres at epsilon 1e-3, 4,096
particles, 10 bisection steps (the single-edge benchmark's options), and the
14-DOF dual arm (64 sphere slots, 96 obstacles) with uneven budgets, whose
bisection runs the 64-slot kernel variant."""
import numpy as np
import pytest

from test_gpu_eizo import bisect_outcomes

pytestmark = pytest.mark.gpu

EPS, DELTA, TAU, STEPS, MARGIN = 1e-3, 0.05, 0.5, 10, 0.01
BAND = 1e-5


def _options(csd, particles, rounds):
    o = csd.EizoOptions()
    o.epsilon, o.delta, o.tau = EPS, DELTA, TAU
    o.bisection_steps, o.num_particles, o.mixing_steps = STEPS, particles, 40
    o.max_hyperplanes_per_iteration = 10
    o.max_iterations = rounds
    o.configuration_margin = MARGIN
    o.verbose = False
    return o


def _domain(csd, plant):
    dom = csd.HPolyhedron()
    dom.MakeBox(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    return dom, np.asarray(dom.A(), np.float32), np.asarray(dom.b(), np.float32)


def _round_samples(gpu, A, b, ls, le, n, round_=0, seed=1337, stream_offset=0):
    """Samples [R][n][dim] of round `round_` of regions (A, b) with segments
    ls/le (dim x R): csd_hit_and_run_segments, region p drawing stream
    p + stream_offset (its index in the EI-ZO batch)."""
    import torch
    from csdecomp_amd import capi
    dim, R = ls.shape
    F = A.shape[0]
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(gpu)  # noqa: E731
    out = torch.empty(R * n * dim, device=gpu)
    capi.hit_and_run_segments(t(np.tile(A.reshape(-1), R)), t(np.tile(b, R)),
                              t(np.full(R, F, np.int32)), F, dim, R, t(ls.T), t(le.T), n, 40,
                              seed, stream_offset, round_, out)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(R, n, dim)


def _check_round(orc, op, samples, M, budget, parts, info, F, what):
    """Stages 1-3 of a round against the oracle's flags of the same samples.
    Returns the number of band samples met."""
    n = samples.shape[0]
    flags = orc.check(op, np.asfortranarray(samples.T), orc.RULE_GPU)
    sd = orc.min_sd(op, np.asfortranarray(samples.T))
    col, band = flags == 0, np.abs(sd) < BAND
    # FP32 flags are the FP64 sign off the band (pins the mirror on this data)
    assert np.all(col[~band] == (sd[~band] < 0)), what
    # round record
    assert info.num_samples_drawn == n, what
    assert info.num_faces_at_start == F, what
    thresh = np.float32(1 - TAU) * np.float32(EPS) * np.float32(M)
    assert info.collisions_threshold == int(thresh), what
    first, first_band = int(col[:M].sum()), int(band[:M].sum())
    assert abs(info.num_collisions_unadaptive - first) <= first_band, (what, info.num_collisions_unadaptive, first)
    assert abs(info.num_collisions_found - int(col.sum())) <= int(band.sum()), what
    certified = info.num_collisions_unadaptive <= thresh
    if abs(first - thresh) > first_band:
        assert certified == (first <= thresh), what
    if certified:
        assert parts.shape[0] == 0 and info.num_collisions_updated == 0, what
        return int(band.sum())
    # particles: the first colliding samples in sample order, up to the budget
    u = parts.shape[0]
    assert info.num_collisions_updated == u == min(info.num_collisions_found, budget), what
    index = {}
    for i in range(n - 1, -1, -1):
        index[samples[i].tobytes()] = i
    idx = np.array([index[p.tobytes()] for p in parts])  # each particle is one of the samples
    assert np.all(np.diff(idx) > 0), f"{what}: particles out of sample order"
    assert np.all(col[idx] | band[idx]), f"{what}: a particle is free in the oracle"
    taken = np.zeros(n, bool)
    taken[idx] = True
    skipped = col & ~band & ~taken
    if u == budget:
        skipped[idx[-1] + 1:] = False  # the budget was full
    assert not skipped.any(), f"{what}: colliding samples {np.nonzero(skipped)[0][:5]} not taken"
    return int(band.sum())


def _check_bisection(orc, op, parts, scol, proj, ls, le, what, limit=None):
    """Stage 4: projections bit-exact; bisected points equal to the oracle's,
    or reachable through a midpoint in the contact band. Returns (exact,
    checked)."""
    k_all = range(parts.shape[0]) if limit is None else range(min(limit, parts.shape[0]))
    exact = 0
    for k in k_all:
        pr, _ = orc.project(parts[k], ls, le)
        assert np.array_equal(proj[k], pr), (what, k)
        ref = orc.bisect(op, parts[k], pr, STEPS)
        if np.array_equal(ref, scol[k]):
            exact += 1
            continue
        outs = bisect_outcomes(orc, op, parts[k], pr, STEPS)
        assert any(np.array_equal(o, scol[k]) for o in outs), (what, k, scol[k], ref)
    return exact, len(k_all)


def _check_faces(orc, region, A0, b0, scol, proj, ls, le, cap, what):
    """Stage 5: the oracle's FP64 greedy placement on the GPU's bisected
    points gives the GPU's faces."""
    F0, dim = A0.shape
    A = np.zeros((cap, dim), np.float32)
    b = np.zeros(cap, np.float32)
    A[:F0], b[:F0] = A0, b0
    added, nf = orc.place_faces(A, b, F0, cap, scol, proj, ls, le, MARGIN, 10)
    gA, gb = np.asarray(region.A()), np.asarray(region.b())
    assert gA.shape[0] == nf, (what, gA.shape[0], nf)
    np.testing.assert_array_equal(gA[:F0], A0)
    np.testing.assert_allclose(gA, A[:nf], rtol=1e-4, atol=1e-6, err_msg=what)
    np.testing.assert_allclose(gb, b[:nf], rtol=1e-4, atol=1e-6, err_msg=what)
    return added


def test_round_sampler_matches_oracle_chains(gpu, orc):
    """csd_hit_and_run_segments against the oracle's spread-start chains
    (orc_hit_and_run_segment): same Philox counters, libm vs device math."""
    from csdecomp_amd import scenes
    import csdecomp_amd as csd
    pl = scenes.arm_scene(1001, 48, 16)
    _, A, b = _domain(csd, pl)
    ls, le = scenes.free_segments(pl, 3, 3001, (0.2, 1.5))
    n = 2048
    g = _round_samples(gpu, A, b, ls, le, n, round_=2)
    for r in range(3):
        ref = orc.hit_and_run_segment(A, b, ls[:, r], le[:, r], n, 40, stream=r, round_=2)
        err = np.abs(g[r] - ref).max(1)
        # rounding differences compound over the 40 steps of a chain
        assert err.max() < 1e-3 and np.median(err) < 2e-5 and np.mean(err < 1e-4) > 0.97, \
            (r, err.max(), np.median(err), np.mean(err < 1e-4))
        assert np.all(g[r] @ A.T <= b + 1e-4)


def test_first_round_parity_arm7(gpu, orc):
    """The single-edge workload's options (epsilon 1e-3, 4,096 particles, 10
    bisection steps, <= 10 faces per round) on 16 of its segments."""
    from csdecomp_amd import scenes
    import csdecomp_amd as csd
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    op = orc.Plant(mp)
    dom, A0, b0 = _domain(csd, pl)
    R, U = 16, 4096
    ls, le = scenes.free_segments(pl, R, 3001, (0.2, 1.5))
    cap = A0.shape[0] + 100
    inf = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), _options(csd, U, 1), dom, R, 10,
                               cap)
    regions, infos = inf.inflateEdges(ls, le, None, 0.0, [])
    M = orc.num_samples_unadaptive(EPS, orc.delta_k(DELTA, 0), TAU)
    assert M == 27948
    samples = _round_samples(gpu, A0, b0, ls, le, max(M, U))
    exact = checked = faces = 0
    for r in range(R):
        info = infos[r].iteration_info
        assert len(info) == 1 == infos[r].iterations
        parts, scol, proj = (np.ascontiguousarray(x.T) for x in inf.last_particles(r))
        _check_round(orc, op, samples[r], M, U, parts, info[0], A0.shape[0], f"region {r}")
        if parts.shape[0] == 0:
            continue
        e, c = _check_bisection(orc, op, parts, scol, proj, ls[:, r], le[:, r], f"region {r}",
                                limit=512)
        exact, checked = exact + e, checked + c
        added = _check_faces(orc, regions[r], A0, b0, scol, proj, ls[:, r], le[:, r], cap,
                             f"region {r}")
        assert added == info[0].num_faces_placed
        faces += added
    assert checked >= 2000 and exact >= 0.95 * checked, (exact, checked)
    assert faces > 0


def test_first_rounds_parity_dual_arm(gpu, orc):
    """The 14-DOF dual arm (64 sphere slots): uneven budgets, two rounds;
    the second round's samples come from the first round's regions. The
    bisection of >= 500 particles runs the 64-slot EI-ZO kernel."""
    from csdecomp_amd import scenes
    import csdecomp_amd as csd
    pl = scenes.dual_arm_scene(1002, 72, 24)
    mp = pl.getMinimalPlant()
    op = orc.Plant(mp)
    dom, A0, b0 = _domain(csd, pl)
    R = 6
    ls, le = scenes.free_segments(pl, R, 3002, (0.05, 2.0), log_uniform=True)
    budgets = [512, 1024, 3000, 700, 2048, 1500]
    cap = A0.shape[0] + 100
    inf = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), _options(csd, 4096, 1), dom,
                               R, 10, cap)
    regions, infos = inf.inflateEdges(ls, le, None, 0.0, budgets)
    M = orc.num_samples_unadaptive(EPS, orc.delta_k(DELTA, 0), TAU)
    samples = _round_samples(gpu, A0, b0, ls, le, M)  # M exceeds every budget
    exact = checked = 0
    for r in range(R):
        info = infos[r].iteration_info
        parts, scol, proj = (np.ascontiguousarray(x.T) for x in inf.last_particles(r))
        _check_round(orc, op, samples[r], M, budgets[r], parts, info[0], A0.shape[0], f"region {r}")
        if parts.shape[0] == 0:
            continue
        e, c = _check_bisection(orc, op, parts, scol, proj, ls[:, r], le[:, r], f"region {r}",
                                limit=200)
        exact, checked = exact + e, checked + c
        _check_faces(orc, regions[r], A0, b0, scol, proj, ls[:, r], le[:, r], cap, f"region {r}")
    assert checked >= 500 and exact >= 0.95 * checked, (exact, checked)

    # round 2 (sampler round 1) in the regions round 1 left: the same
    # stages against the oracle, on the polytopes the first run returned
    inf2 = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), _options(csd, 4096, 2), dom,
                                R, 10, cap)
    regions2, infos2 = inf2.inflateEdges(ls, le, None, 0.0, budgets)
    M1 = orc.num_samples_unadaptive(EPS, orc.delta_k(DELTA, 1), TAU)
    seen = 0
    for r in range(R):
        if infos[r].converged or infos[r].face_capacity_reached:
            continue
        assert infos2[r].iterations == 2
        Ar, br = np.asarray(regions[r].A(), np.float32), np.asarray(regions[r].b(), np.float32)
        # round 1 of the two-round run is the one-round run
        np.testing.assert_array_equal(np.asarray(regions2[r].A())[:Ar.shape[0]], Ar)
        n1 = max(M1, budgets[r])
        g = _round_samples(gpu, Ar, br, ls[:, r:r + 1], le[:, r:r + 1], n1, round_=1,
                           stream_offset=r)[0]
        parts, scol, proj = (np.ascontiguousarray(x.T) for x in inf2.last_particles(r))
        info = infos2[r].iteration_info[1]
        _check_round(orc, op, g, M1, budgets[r], parts, info, Ar.shape[0], f"region {r} round 2")
        if parts.shape[0]:
            _check_bisection(orc, op, parts, scol, proj, ls[:, r], le[:, r],
                             f"region {r} round 2", limit=64)
            _check_faces(orc, regions2[r], Ar, br, scol, proj, ls[:, r], le[:, r], cap,
                         f"region {r} round 2")
        seen += 1
        if seen == 2:
            break
    assert seen >= 1


def _fresh_fraction(gpu, mp, P, ls, le, n=20000, stream=0x30000000):
    """Collision fraction (GPU checker) of n fresh samples of the
    certificate's own distribution: 40-step chains spread along the segment,
    random streams no round uses."""
    import csdecomp_amd as csd
    A, b = np.asarray(P.A(), np.float32), np.asarray(P.b(), np.float32)
    s = _round_samples(gpu, A, b, ls[:, None], le[:, None], n, round_=0, seed=4242,
                       stream_offset=stream)[0]
    return 1.0 - float(np.mean(csd.CheckCollisionFreeCuda(np.asfortranarray(s.T), mp)))


def _uniform_fraction(mp, P, ls, le, n=20000):
    """Collision fraction of n samples of 200-step chains from the segment
    midpoint: closer to uniform over the region than 40 steps."""
    import csdecomp_amd as csd
    c = ((ls + le) / 2).astype(np.float32)
    s = csd.UniformSampleInHPolyhedraCuda([P], c[:, None], n, 200, seed=77)[0]
    return 1.0 - float(np.mean(csd.CheckCollisionFreeCuda(s, mp)))


def _region_properties(gpu, mp, regions, infos, ls, le, eps, what):
    """Both endpoints inside every region; certified regions pass a fresh
    test of their certificate's distribution; the round records add up."""
    fractions = []
    for r, (P, info) in enumerate(zip(regions, infos)):
        assert P.PointInSet(ls[:, r], 1e-5) and P.PointInSet(le[:, r], 1e-5), (what, r)
        rounds = info.iteration_info
        assert len(rounds) == info.iterations, (what, r)
        assert sum(x.num_faces_placed for x in rounds) == info.faces_added, (what, r)
        assert sum(x.num_samples_drawn for x in rounds) == info.samples_drawn, (what, r)
        if info.converged:
            last = rounds[-1]
            assert last.num_collisions_unadaptive <= last.collisions_threshold + 1, (what, r)
            f = _fresh_fraction(gpu, mp, P, ls[:, r], le[:, r])
            assert f < 2.5 * eps, (what, r, f)
            fractions.append(f)
        else:
            assert info.face_capacity_reached or info.iterations == len(rounds), (what, r)
    return fractions


def _graph_equals_host(csd, mp, pl, opt, dom, ls, le, budgets, cap, regions, infos, what):
    inf = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), opt, dom, ls.shape[1], 10, cap)
    inf.set_graph(4)
    g_regions, g_infos = inf.inflateEdges(ls, le, None, 0.0, budgets)
    for r in range(ls.shape[1]):
        np.testing.assert_array_equal(g_regions[r].A(), regions[r].A(), err_msg=f"{what} {r}")
        np.testing.assert_array_equal(g_regions[r].b(), regions[r].b(), err_msg=f"{what} {r}")
        assert (g_infos[r].iterations, g_infos[r].converged, g_infos[r].samples_drawn) == \
            (infos[r].iterations, infos[r].converged, infos[r].samples_drawn), (what, r)
        assert [x.num_faces_placed for x in g_infos[r].iteration_info] == \
            [x.num_faces_placed for x in infos[r].iteration_info], (what, r)


def test_single_edge_workload_properties(gpu, orc):
    """The single-edge workload's own options (epsilon 1e-3, delta 0.05,
    tau 0.5, 4,096 particles, 10 bisection steps, <= 10 faces per round,
    <= 100 faces beyond the domain) on 16 of its segments, to the end."""
    from csdecomp_amd import scenes
    import csdecomp_amd as csd
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    dom, A0, _ = _domain(csd, pl)
    ls, le = scenes.free_segments(pl, 16, 3001, (0.2, 1.5))
    cap = A0.shape[0] + 100
    opt = _options(csd, 4096, 500)
    inf = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), opt, dom, 16, 10, cap)
    regions, infos = inf.inflateEdges(ls, le, None, 0.0, [])
    fr = _region_properties(gpu, mp, regions, infos, ls, le, EPS, "arm7")
    assert len(fr) >= 10  # most regions certify at this face budget
    # the uniform check is not the certificate's distribution; in 7 dims the
    # 40-step chains reach the region well enough that it holds too
    for r, (P, info) in enumerate(zip(regions, infos)):
        if info.converged:
            assert _uniform_fraction(mp, P, ls[:, r], le[:, r]) < 3 * EPS, r
    _graph_equals_host(csd, mp, pl, opt, dom, ls, le, [], cap, regions, infos, "arm7 graph")


def test_batched_dual_arm_workload_properties(gpu, orc):
    """The batched workload: 14-DOF dual arm, segments log-uniform in
    [0.05, 2] rad, budgets log-uniform in [512, 16384], epsilon 1e-3, <= 100
    faces beyond the domain; 24 regions in one batch, host-driven and
    graph-captured rounds."""
    from csdecomp_amd import scenes
    import csdecomp_amd as csd
    pl = scenes.dual_arm_scene(1002, 72, 24)
    mp = pl.getMinimalPlant()
    dom, A0, _ = _domain(csd, pl)
    R = 24
    ls, le = scenes.free_segments(pl, R, 3002, (0.05, 2.0), log_uniform=True)
    budgets = scenes.log_uniform_budgets(3003, R)
    assert min(budgets) < 2000 and max(budgets) > 5000  # uneven
    cap = A0.shape[0] + 100
    opt = _options(csd, 16384, 500)
    inf = csd.CudaEdgeInflator(mp, list(pl.getRobotGeometryIds()), opt, dom, R, 10, cap)
    regions, infos = inf.inflateEdges(ls, le, None, 0.0, budgets)
    fr = _region_properties(gpu, mp, regions, infos, ls, le, EPS, "dual arm")
    assert len(fr) >= 1, [i.converged for i in infos]
    assert all(i.converged or i.face_capacity_reached for i in infos)
    _graph_equals_host(csd, mp, pl, opt, dom, ls, le, budgets, cap, regions, infos,
                       "dual arm graph")


def test_ellipsoid_metric_face_placement(gpu, orc):
    """Faces placed against an ellipsoid metric (csd_eizo_set_metric): the
    GPU's FP64 placement on its own bisected points equals the oracle's
    orc_place_faces_metric; the identity metric gives the Euclidean run bit
    for bit; full runs with the metric keep the segments inside."""
    from csdecomp_amd import scenes
    import csdecomp_amd as csd
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    dom, A0, b0 = _domain(csd, pl)
    R = 8
    ls, le = scenes.free_segments(pl, R, 3001, (0.2, 1.5))
    cap = A0.shape[0] + 100
    rng = np.random.default_rng(5)
    G = rng.normal(size=(7, 7))
    M = (G @ G.T / 7 + np.diag([4.0, 4.0, 2.0, 2.0, 1.0, 1.0, 0.5])).astype(np.float32)
    M = (M + M.T) / 2
    rg = list(pl.getRobotGeometryIds())
    inf = csd.CudaEdgeInflator(mp, rg, _options(csd, 4096, 1), dom, R, 10, cap)
    inf.set_metric(M)
    regions, infos = inf.inflateEdges(ls, le, None, 0.0, [])
    placed = 0
    for r in range(R):
        parts, scol, proj = (np.ascontiguousarray(x.T) for x in inf.last_particles(r))
        if parts.shape[0] == 0:
            continue
        A = np.zeros((cap, 7), np.float32)
        b = np.zeros(cap, np.float32)
        A[:14], b[:14] = A0, b0
        added, nf = orc.place_faces(A, b, 14, cap, scol, proj, ls[:, r], le[:, r], MARGIN, 10,
                                    metric=M.astype(np.float64))
        gA, gb = np.asarray(regions[r].A()), np.asarray(regions[r].b())
        assert gA.shape[0] == nf, r
        np.testing.assert_allclose(gA, A[:nf], rtol=1e-4, atol=1e-6, err_msg=f"region {r}")
        np.testing.assert_allclose(gb, b[:nf], rtol=1e-4, atol=1e-6, err_msg=f"region {r}")
        # the metric changed the faces: they are not the Euclidean ones
        Ae = np.zeros((cap, 7), np.float32)
        be = np.zeros(cap, np.float32)
        Ae[:14], be[:14] = A0, b0
        orc.place_faces(Ae, be, 14, cap, scol, proj, ls[:, r], le[:, r], MARGIN, 10)
        assert not np.allclose(Ae[14:nf], gA[14:], atol=1e-3)
        placed += added
    assert placed > 0
    # identity metric == Euclidean, bit for bit; None restores the Euclidean metric
    inf.set_metric(np.eye(7, dtype=np.float32))
    r_id, _ = inf.inflateEdges(ls, le, None, 0.0, [])
    inf.set_metric(None)
    r_eu, _ = inf.inflateEdges(ls, le, None, 0.0, [])
    for r in range(R):
        np.testing.assert_array_equal(r_id[r].A(), r_eu[r].A())
        np.testing.assert_array_equal(r_id[r].b(), r_eu[r].b())
    with pytest.raises(ValueError):
        inf.set_metric(np.diag([1, 1, 1, -1, 1, 1, 1]).astype(np.float32))
    with pytest.raises(ValueError):
        bad = np.eye(7, dtype=np.float32)
        bad[0, 1] = 0.5
        inf.set_metric(bad)
    # a full run against the metric: segments inside, certified regions pass
    # a fresh test of their distribution
    inf2 = csd.CudaEdgeInflator(mp, rg, _options(csd, 4096, 500), dom, R, 10, cap)
    inf2.set_metric(M)
    regions2, infos2 = inf2.inflateEdges(ls, le, None, 0.0, [])
    fr = _region_properties(gpu, mp, regions2, infos2, ls, le, EPS, "metric")
    assert len(fr) >= 1


def test_trained_surrogate_gain(gpu, orc):
    """What the guidance and the trained network gain, on the single-edge
    workload's options with room for 300 faces (every region certifies in
    all modes); the exact checker judges every sample that counts.
    - Guided rounds (an exact pilot of 8,192 samples, 256 surrogate-ranked
      particles bisected) certify plain EI-ZO's fraction with fewer exact
      collision checks of samples.
    - The trained network's own gain is in the pruned pilot: the pilot
      samples it predicts free (above 0.1, i.e. 2 cm) skip the exact check.
      Trained, it prunes a large share of the pilot; a freshly initialised
      network of the same shape prunes far less. Certification is unchanged:
      the unadaptive test is never pruned."""
    from csdecomp_amd import capi, scenes
    import csdecomp_amd as csd
    from test_gpu_eizo import _train_surrogate
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    dom, A0, _ = _domain(csd, pl)
    R = 48
    ls, le = scenes.free_segments(pl, R, 3001, (0.2, 1.5))
    cap = A0.shape[0] + 300
    rg = list(pl.getRobotGeometryIds())
    trained = _train_surrogate(gpu, pl, ls, le, steps=6000, n=2_000_000, seed=7)
    lo = np.asarray(pl.getPositionLowerLimits(), np.float32)
    hi = np.asarray(pl.getPositionUpperLimits(), np.float32)
    untrained = capi.Mlp(7, lo, hi, 7)
    out = {}
    runs = (("plain", None, None), ("guided", trained, None), ("pruned_trained", trained, 0.1),
            ("pruned_untrained", untrained, 0.1))
    for mode, model, prune in runs:
        opt = _options(csd, 4096, 500)
        opt.track_iteration_information = True
        inf = csd.CudaEdgeInflator(mp, rg, opt, dom, R, 10, cap)
        if model is not None:
            inf.set_guidance(pilot_samples=8192, pilot_factor=10.0, surrogate=model,
                             keep_particles=256, line_points=4, prune_below=prune)
        regions, infos = inf.inflateEdges(ls, le, None, 0.0, [])
        for r in range(R):
            assert regions[r].PointInSet(ls[:, r], 1e-5) and regions[r].PointInSet(le[:, r], 1e-5)
        rounds = [x for i in infos for x in i.iteration_info]
        pilot = sum(x.num_pilot_samples for x in rounds)
        checked = sum(x.num_pilot_checked for x in rounds)
        out[mode] = (sum(i.converged for i in infos), sum(i.samples_drawn for i in infos),
                     checked / max(pilot, 1))
    (cp, sp, _), (cg, sg, fg) = out["plain"], out["guided"]
    assert cg >= cp >= R - 1, out
    assert sg < 0.9 * sp, out  # fewer exact sample checks
    assert fg == 1.0, out  # unpruned pilot: every sample checked
    ct, _, ft = out["pruned_trained"]
    cu, _, fu = out["pruned_untrained"]
    assert ct >= cp - 1 and cu >= cp - 1, out
    assert ft < 0.8 and ft < fu - 0.15, out  # the training is what prunes the pilot
