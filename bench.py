"""Benchmark: collision-checked configurations per second on one MI355X.

Workload (default): 7-DOF serial revolute arm, 8 links x 4 spheres, 16
world-fixed oriented boxes (784 collision pairs), 65,536 joint vectors
uniform in [-pi, pi]^7 from a fixed seed, already resident in HBM as
structure-of-arrays. A step is one fused FK + collision-check launch over
the whole batch, producing one uint8 flag per configuration.

Prints one JSON line. `roofline` is computed from the algorithmic FLOPs per
configuration (132 per revolute joint + 21 per moving sphere + 36 per
sphere-box pair + 9 per sphere-sphere pair = 20,664 for this plant) over the
kernel's average launch time measured with HIP events on its stream.
`cpu_baseline` times the CPU restatement of the reference's CPU checker
(early exit, OpenMP over configurations) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X, vendor figure
HBM_PEAK_GBS = 8000.0


def algorithmic_flops_per_config(mp, plant):
    """132*J_rev + 21*S_mov + 36*P_sb + 9*P_ss + 60*P_bb (spheres/boxes)."""
    geoms = mp.scene_collision_geometries
    tree = mp.kin_tree
    j_rev = sum(1 for j in tree.joints if int(j.type) == 1)
    robot = set(plant.getRobotGeometryIds())
    s_mov = sum(1 for g in robot if int(geoms[g].type) == 1)
    pairs = np.asarray(mp.collision_pairs_flat).reshape(-1, 2)
    p_sb = p_ss = p_bb = 0
    for a, b in pairs:
        ta, tb = int(geoms[a].type), int(geoms[b].type)
        if ta == tb == 1:
            p_ss += 1
        elif ta == tb == 0:
            p_bb += 1
        else:
            p_sb += 1
    return 132 * j_rev + 21 * s_mov + 36 * p_sb + 9 * p_ss + 60 * p_bb


def plant_dof(plant):
    return len(plant.getPositionLowerLimits())


def cpu_baseline(mp, q, seconds):
    """CPU restatement of checkCollisionFreeVoxels with no voxels (early exit,
    silent), OpenMP over configurations on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    op = oracle.Plant(mp)
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1
    n = q.shape[1]
    oracle.check(op, q[:, :1024], oracle.RULE_CPU, threads, fast=True)  # warm up
    done, t0 = 0, time.perf_counter()
    while True:
        oracle.check(op, q, oracle.RULE_CPU, threads, fast=True)
        done += n
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": done / el, "unit": "configs/s", "cores": threads, "kind": "port",
            "sample": f"{done} configs ({done // n} passes over the {n}-config batch), {el:.1f} s, "
                      f"CPU rule of collision_checker.cpp with early exit, speed build "
                      f"(-O3, AVX2, FMA contraction)"}


def eizo_cpu_baseline(plant, starts, ends, seconds=12.0, face_budget=100):
    """CPU restatement of single-edge EI-ZO (the oracle's orc_eizo: the
    rounds of cuda_edge_inflation_zero_order.cu:94-339 with the GPU leg's
    options) over the leg's first segments, on every host core: one region
    per OpenMP thread (orc_eizo_batch, built for speed), batches of one
    region per thread until `seconds` have passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import csdecomp_amd as csd
    op = oracle.Plant(plant.getMinimalPlant())
    dom = csd.HPolyhedron()
    dom.MakeBox(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    A, b = np.asarray(dom.A(), np.float32), np.asarray(dom.b(), np.float32)
    # margin, tau, delta, epsilon, bisection steps, particles, faces per round,
    # rounds, mixing steps, seed
    opts = (0.01, 0.5, 0.05, 1e-3, 10, 4096, 10, 500, 40, 1337)
    rg = list(plant.getRobotGeometryIds())
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1
    done = certified = 0
    t0 = time.perf_counter()
    el = 0.0
    while done < starts.shape[1] and el < seconds:
        k = min(threads, starts.shape[1] - done)
        _, _, _, st, infos = oracle.eizo_batch(
            op, rg, A, b, np.ascontiguousarray(starts[:, done:done + k]),
            np.ascontiguousarray(ends[:, done:done + k]), opts, cap=A.shape[0] + face_budget,
            spread_starts=True, threads=threads, fast=True)
        # a region that ran out of faces (status -1) ends uncertified, as on the GPU
        certified += sum(int(i.converged) for i, s_ in zip(infos, st) if s_ == 0)
        done += k
        el = time.perf_counter() - t0
    return {"value": done / el, "unit": "regions/s", "cores": threads, "kind": "port",
            "certified_fraction": certified / done,
            "sample": f"{done} regions (the first segments of the leg), {el:.1f} s; one region "
                      f"per thread, speed build of the restatement (FMA contraction)"}


def eizo_region_flops(info, dim, mixing, steps, f_cfg):
    """Algorithmic FLOPs of one region, summed over its rounds from the
    kernels' own counts (EizoInfo.iteration_info):
      N_k mix (2 F d + 3 F + 2 d)   hit-and-run: two dot products, the chord
                                    bound and the step per face and step
    + (N_k + u steps) F_cfg         collision checks of samples and bisection
    + faces (2 d + u 2 d)           FP64 face placement and redundancy test
    with N_k the samples of the round (test + pilot), F the region's faces
    when the round began, u the particles bisected."""
    tot = 0
    for it in info.iteration_info:
        nk = it.num_samples_drawn + it.num_pilot_samples
        F, u, faces = it.num_faces_at_start, it.num_collisions_updated, it.num_faces_placed
        tot += nk * mixing * (2 * F * dim + 3 * F + 2 * dim)
        tot += (nk + u * steps) * f_cfg
        tot += faces * (2 * dim + u * 2 * dim)
    return tot


def eizo_leg(name, plant, starts, ends, budgets, particles, modes, model, face_budget=100,
             train_s=0.0, pilot=8192, keep=256):
    """Regions/s of batched EI-ZO (epsilon 1e-3, delta 0.05, tau 0.5, 10
    bisection steps, 40 mixing steps, <= 10 faces per round, <= face_budget
    faces beyond the domain) in each mode. Guided modes count the
    surrogate's labelling and training time (train_s) in their regions/s and
    also report the inflation alone. "guided_untrained" runs the same guided
    rounds with a freshly initialised network (no training time): the
    difference to "guided" is what the training itself gains."""
    import torch
    import csdecomp_amd as csd
    mp = plant.getMinimalPlant()
    dom = csd.HPolyhedron()
    dom.MakeBox(plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    nd = dom.A().shape[0]
    opt = csd.EizoOptions()
    opt.epsilon, opt.delta, opt.tau = 1e-3, 0.05, 0.5
    opt.bisection_steps, opt.num_particles, opt.mixing_steps = 10, particles, 40
    opt.max_hyperplanes_per_iteration = 10
    opt.verbose = False
    R = starts.shape[1]
    f_cfg = algorithmic_flops_per_config(mp, plant)
    out = {"segments": R, "dof": int(starts.shape[0])}
    faces = {}
    wide_from = None
    for mode in modes:
        inf = csd.CudaEdgeInflator(mp, list(plant.getRobotGeometryIds()), opt, dom, R, 10, nd + face_budget)
        # sampler tile: timed once per leg on synthetic polytopes (outside the timed region)
        if wide_from is None:
            wide_from = inf.tune_sampler()
            out["sampler_wide_from_faces"] = wide_from
        else:
            inf.set_sampler_tile(wide_from)
        if mode in ("plain_graph", "guided_graph"):
            inf.set_graph(1)  # one graph-captured round per launch, status read each round
        if mode in ("guided", "guided_graph", "guided_fallback"):
            inf.set_guidance(pilot_samples=pilot, pilot_factor=10.0, surrogate=model,
                             keep_particles=keep, line_points=4,
                             plain_fallback=mode == "guided_fallback")
        if mode in ("guided_pruned", "guided_pruned_graph"):
            inf.set_guidance(pilot_samples=pilot, pilot_factor=10.0, surrogate=model,
                             keep_particles=keep, line_points=4, prune_below=PRUNE_BELOW)
            if mode == "guided_pruned_graph":
                inf.set_graph(1)
        if mode in ("guided_untrained", "guided_pruned_untrained"):
            from csdecomp_amd import capi
            lo = np.asarray(plant.getPositionLowerLimits(), np.float32)
            hi = np.asarray(plant.getPositionUpperLimits(), np.float32)
            inf.set_guidance(pilot_samples=pilot, pilot_factor=10.0,
                             surrogate=capi.Mlp(len(lo), lo, hi, 7), keep_particles=keep,
                             line_points=4,
                             prune_below=PRUNE_BELOW if mode == "guided_pruned_untrained" else None)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        regions, infos = inf.inflateEdges(starts, ends, None, 0.0, budgets)
        torch.cuda.synchronize()
        infl = time.perf_counter() - t0
        trained = mode in ("guided", "guided_graph", "guided_fallback", "guided_pruned",
                           "guided_pruned_graph")
        el = infl + (train_s if trained else 0.0)
        conv = np.array([i.converged for i in infos])
        faces[mode] = [(r.A(), r.b()) for r in regions]
        flops = sum(eizo_region_flops(i, starts.shape[0], opt.mixing_steps, opt.bisection_steps,
                                      f_cfg) for i in infos)
        out[mode] = {"seconds": round(el, 3), "regions_per_s": round(R / el, 2),
                     "inflation_seconds": round(infl, 3),
                     "inflation_regions_per_s": round(R / infl, 2),
                     "certified_per_s": round(float(conv.sum()) / el, 2),
                     "certified_fraction": round(float(conv.mean()), 4),
                     "rounds_mean": round(float(np.mean([i.iterations for i in infos])), 2),
                     "faces_mean": round(float(np.mean([r.A().shape[0] - nd for r in regions])), 1),
                     "samples": int(sum(i.samples_drawn for i in infos)),
                     "roofline": {"bound": "valu", "achieved": flops / infl / 1e12,
                                  "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": flops / infl / 1e12 / FP32_VECTOR_PEAK_TFLOPS,
                                  "flops": flops, "flops_per_region": flops / R,
                                  "f_cfg": f_cfg}}
        if mode == "guided_fallback":
            out[mode]["fallback_regions"] = int(sum(i.plain_fallback for i in infos))
        rounds = [x for i in infos for x in i.iteration_info]
        pilot_n = sum(x.num_pilot_samples for x in rounds)
        if pilot_n:
            out[mode]["pilot_checked_fraction"] = round(
                sum(x.num_pilot_checked for x in rounds) / pilot_n, 4)
        print(f"eizo {name} R={R} faces<={face_budget} {mode}: {out[mode]['regions_per_s']} "
              f"regions/s, certified {out[mode]['certified_fraction']}", file=sys.stderr,
              flush=True)
    for host, dev in (("plain", "plain_graph"), ("guided", "guided_graph"),
                      ("guided_pruned", "guided_pruned_graph")):
        if host in faces and dev in faces:
            out[dev + "_identical"] = all(
                np.array_equal(a[0], g[0]) and np.array_equal(a[1], g[1])
                for a, g in zip(faces[host], faces[dev]))
    if "plain" in out and "guided" in out:
        out["guided_vs_plain_regions_per_s"] = round(out["guided"]["regions_per_s"] / out["plain"]["regions_per_s"], 3)
        out["guided_vs_plain_samples"] = round(out["guided"]["samples"] / out["plain"]["samples"], 3)
        out["guided_vs_plain_certified_per_s"] = round(
            out["guided"]["certified_per_s"] / max(out["plain"]["certified_per_s"], 1e-9), 3)
    if "guided" in out and "guided_untrained" in out:
        g, u = out["guided"], out["guided_untrained"]
        out["trained_vs_untrained_inflation_regions_per_s"] = round(
            g["inflation_regions_per_s"] / u["inflation_regions_per_s"], 3)
        out["trained_vs_untrained_samples"] = round(g["samples"] / u["samples"], 3)
    if "guided_pruned" in out and "guided_pruned_untrained" in out:
        g, u = out["guided_pruned"], out["guided_pruned_untrained"]
        out["pruned_trained_vs_untrained_inflation_regions_per_s"] = round(
            g["inflation_regions_per_s"] / u["inflation_regions_per_s"], 3)
    if "plain" in out and "guided_pruned" in out:
        out["pruned_vs_plain_certified_per_s"] = round(
            out["guided_pruned"]["certified_per_s"] / max(out["plain"]["certified_per_s"], 1e-9), 3)
    return out


MLP_FLOPS_PER_ROW = 2 * (256 * 7 + 3 * 256 * 256 + 256)  # 7 inputs, 4 hidden layers, head
# pruned pilots: samples the surrogate predicts above this value skip the
# exact check (its output is the signed distance clipped to +-0.2 m and
# scaled to +-1, so 0.1 is 2 cm clear)
PRUNE_BELOW = 0.1
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense


def mlp_benchmark(rows=1 << 20, batch=8192, reps=20):
    """Surrogate throughput: fused MFMA inference on `rows` configurations
    (7 inputs) and one training step (forward, backward, Adam) at `batch`,
    both timed with HIP events on the launch stream."""
    import torch
    from csdecomp_amd import capi
    lo, hi = np.full(7, -np.pi, np.float32), np.full(7, np.pi, np.float32)
    m = capi.Mlp(7, lo, hi, 3)
    q = torch.rand(rows, 7, device="cuda") * 6.28 - 3.14
    out = torch.empty(rows, device="cuda")
    t = torch.rand(batch, device="cuda") * 2 - 1
    s = torch.cuda.current_stream()
    for _ in range(3):
        m.predict(q, rows, 7, 1, out, s)
        m.train_step(q[:batch], t, batch, 1e-3, None, s)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        m.predict(q, rows, 7, 1, out, s)
    b.record(s)
    torch.cuda.synchronize()
    inf_ms = a.elapsed_time(b) / reps
    a.record(s)
    for _ in range(reps):
        m.train_step(q[:batch], t, batch, 1e-3, None, s)
    b.record(s)
    torch.cuda.synchronize()
    train_ms = a.elapsed_time(b) / reps
    tflops = rows * MLP_FLOPS_PER_ROW / (inf_ms * 1e-3) / 1e12
    return {"inference_rows": rows, "inference_ms": round(inf_ms, 4),
            "inference_rows_per_s": round(rows / (inf_ms * 1e-3)),
            "inference_tflops": round(tflops, 1), "bf16_peak_tflops": BF16_MFMA_PEAK_TFLOPS,
            "inference_frac": round(tflops / BF16_MFMA_PEAK_TFLOPS, 4),
            "train_batch": batch, "train_step_ms": round(train_ms, 4),
            "train_tflops": round(3 * batch * MLP_FLOPS_PER_ROW / (train_ms * 1e-3) / 1e12, 1)}


def edge_point_counts(dp, q, edges, step, chunk=1 << 24):
    """Work of the checkEdgeCollisionFree discretisation (n = ceil(|q2 - q1| /
    step) points at t = k / n) over `edges` [E][2] of nodes q [N][dim]
    (device tensors), measured with the checker on the materialised points
    outside any timed region: the points up to and including each edge's
    first colliding one (all n for a free edge), which is the least work a
    sequential check can do, and the lane evaluations of the edge kernel
    (64 points per pass, passes until one holds a collision). Also returns
    how many edges the point checks call free."""
    import torch
    from csdecomp_amd import capi
    E, dim = edges.shape[0], q.shape[1]
    q1, q2 = q[edges[:, 0].long()], q[edges[:, 1].long()]
    npts = torch.ceil(torch.linalg.vector_norm(q2 - q1, dim=1) / step).to(torch.int64)
    needed = evaluated = free = 0
    e0 = 0
    csum = torch.cumsum(npts, 0)
    while e0 < E:
        base = int(csum[e0 - 1].item()) if e0 else 0
        e1 = int(torch.searchsorted(csum, torch.tensor(base + chunk, device=q.device)).item())
        e1 = max(e1, e0 + 1)
        e1 = min(e1, E)
        n = npts[e0:e1]
        P = int(n.sum().item())
        if P == 0:
            free += e1 - e0
            e0 = e1
            continue
        eid = torch.repeat_interleave(torch.arange(e1 - e0, device=q.device), n)
        off = torch.cumsum(n, 0) - n
        k = torch.arange(P, device=q.device) - off[eid]
        t = (k.double() / n[eid].double()).float()[:, None]
        pts = (t * q2[e0:e1][eid] + (1 - t) * q1[e0:e1][eid]).contiguous()
        flags = torch.empty(P, dtype=torch.uint8, device=q.device)
        capi.check_collision_free(dp, pts, P, dim, 1, flags)
        first = n.clone()
        col = flags == 0
        first.scatter_reduce_(0, eid[col], k[col], reduce="amin")
        hit = first < n
        needed += int(torch.where(hit, first + 1, n).sum().item())
        passes = torch.where(hit, first // 64 + 1, (n + 63) // 64)
        evaluated += int((passes * 64).sum().item())
        free += int((~hit).sum().item())
        e0 = e1
    return needed, evaluated, free


def occupancy_tests(dp, plant, q, center, size, res):
    """Sphere-voxel tests of the occupancy kernel: for every configuration
    and robot sphere, the candidate voxels of its bounding box (the kernel's
    own bounds: extent r * 1.0001 + voxel radius + 1e-6, one voxel of slack
    on each side, clipped to the grid), from FK on the GPU."""
    import torch
    from csdecomp_amd import capi
    mp = plant.getMinimalPlant()
    geoms = mp.scene_collision_geometries
    n = q.shape[0]
    L = mp.kin_tree.num_links
    T = torch.empty(n, L, 16, device=q.device)
    capi.forward_kinematics(dp, q, n, q.shape[1], 1, T)
    T = T.view(n, L, 4, 4).transpose(2, 3)  # column-major to row-major
    vr = float(np.sqrt(3.0) * res)
    two = 2.0 * res
    nn = [int(np.ceil(size[k] / two)) for k in range(3)]
    tests = 0
    for gid in plant.getRobotGeometryIds():
        g = geoms[gid]
        if int(g.type) != 1:
            raise ValueError("occupancy test count covers sphere geometries only")
        X = torch.as_tensor(np.asarray(g.X_L_B, np.float32), device=q.device)
        p = (T[:, g.link_index] @ X)[:, :3, 3]
        ext = float(np.float32(g.dimensions[0]) * np.float32(1.0001)) + vr + 1e-6
        cnt = torch.ones(n, dtype=torch.int64, device=q.device)
        for k in range(3):
            origin = center[k] - 0.5 * size[k] + res
            lo = torch.clamp(torch.floor((p[:, k] - ext - origin) / two).long() - 1, min=0)
            hi = torch.clamp(torch.ceil((p[:, k] + ext - origin) / two).long() + 1, max=nn[k] - 1)
            cnt *= torch.clamp(hi - lo + 1, min=0)
        tests += int(cnt.sum().item())
    return tests


def roadmap_benchmark(nodes=200_000, k=10, step=0.01):
    """Roadmap construction on the 7-DOF arm with 48 boxes + 16 spheres:
    `nodes` configurations uniform in the limits, brute-force k-NN (squared L2
    in joint space), every undirected k-NN edge validated at `step` rad with
    the checkEdgeCollisionFree discretisation, and the robot-swept occupancy
    of all nodes on a 64^3 grid over [-1, 1]^3 m. Timed with HIP events."""
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    dp = capi.DevicePlant(mp)
    lo = np.asarray(pl.getPositionLowerLimits(), np.float32)
    hi = np.asarray(pl.getPositionUpperLimits(), np.float32)
    q = torch.from_numpy(np.random.default_rng(2003).uniform(lo, hi, (nodes, 7)).astype(np.float32)).cuda()
    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    idx = torch.empty(nodes, k, dtype=torch.int32, device="cuda")
    ev[0].record(s)
    capi.knn(q, nodes, 7, k, idx, None, s)
    ev[1].record(s)
    # undirected, de-duplicated edge list (i < j)
    src = torch.arange(nodes, device="cuda", dtype=torch.int64).repeat_interleave(k)
    dst = idx.reshape(-1).to(torch.int64)
    a, b = torch.minimum(src, dst), torch.maximum(src, dst)
    keys = torch.unique(a * nodes + b)
    edges = torch.stack([keys // nodes, keys % nodes], dim=1).to(torch.int32).contiguous()
    E = edges.shape[0]
    free = torch.empty(E, dtype=torch.uint8, device="cuda")
    length = torch.linalg.vector_norm(q[edges[:, 0].long()] - q[edges[:, 1].long()], dim=1)
    points = int(torch.ceil(length / step).sum().item())
    torch.cuda.synchronize()
    ev[2].record(s)
    capi.validate_edges(dp, q, nodes, edges, E, step, 0, free, s)
    ev[3].record(s)
    grid = torch.zeros(64, 64, 64, dtype=torch.uint8, device="cuda")
    ev[4].record(s)
    capi.occupancy(dp, q, nodes, 7, 1, list(pl.getRobotGeometryIds()), [0, 0, 0], [2, 2, 2], 1.0 / 64,
                   grid, s)
    ev[5].record(s)
    torch.cuda.synchronize()
    knn_ms, edge_ms, occ_ms = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3]), ev[4].elapsed_time(ev[5])
    # algorithmic work, counted outside the timed region
    f_cfg = algorithmic_flops_per_config(mp, pl)
    needed, evaluated, free_pts = edge_point_counts(dp, q, edges, step)
    j_rev = sum(1 for j in mp.kin_tree.joints if int(j.type) == 1)
    s_mov = len(pl.getRobotGeometryIds())
    tests = occupancy_tests(dp, pl, q, [0, 0, 0], [2, 2, 2], 1.0 / 64)
    occ_flops = nodes * (132 * j_rev + 21 * s_mov) + 9 * tests
    edge_tf = needed * f_cfg / (edge_ms * 1e-3) / 1e12
    occ_tf = occ_flops / (occ_ms * 1e-3) / 1e12
    return {"nodes": nodes, "k": k, "knn_ms": round(knn_ms, 3), "edges": E,
            "edge_points_max": points, "edges_free": int(free.sum().item()),
            "edges_free_by_point_checks": free_pts,
            "edge_validation_ms": round(edge_ms, 3),
            "edges_per_s": round(E / (edge_ms * 1e-3)),
            "edge_points_needed": needed, "edge_lane_evaluations": evaluated,
            "edge_points_per_s": round(needed / (edge_ms * 1e-3)),
            "edge_roofline": {"bound": "valu", "achieved": edge_tf, "peak": FP32_VECTOR_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": edge_tf / FP32_VECTOR_PEAK_TFLOPS,
                              "flops_per_point": f_cfg,
                              "note": "points up to each edge's first collision x F_cfg"},
            "occupancy_voxels": int((grid > 0).sum().item()), "occupancy_ms": round(occ_ms, 3),
            "occupancy_configs_per_s": round(nodes / (occ_ms * 1e-3)),
            "occupancy_voxel_tests": tests,
            "occupancy_roofline": {"bound": "valu", "achieved": occ_tf,
                                   "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                                   "frac": occ_tf / FP32_VECTOR_PEAK_TFLOPS,
                                   "note": "FK 132 J_rev + 21 per sphere + 9 per sphere-voxel test"}}


def edit_regions_benchmark(regions=2048, per=32, steps=9, voxels=64, reps=20):
    """EditRegions on the 7-DOF arm with 48 boxes + 16 spheres and `voxels`
    voxel spheres (radius 0.02 m): `regions` boxes around free segments
    (0.2-0.8 rad) widened by 0.4 rad, and the first `per` colliding of 256
    configurations sampled uniformly in each box. One call of
    csd_edit_regions = projection + `steps` bisection steps per collision,
    then the faces of every region. Timed with HIP events on the launch
    stream; each call first restores the regions (3 small device copies)."""
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1001, 48, 16)
    mp = pl.getMinimalPlant()
    dp = capi.DevicePlant(mp)
    lo = np.asarray(pl.getPositionLowerLimits(), np.float32)
    hi = np.asarray(pl.getPositionUpperLimits(), np.float32)
    ls, le = scenes.free_segments(pl, regions, 4001, (0.2, 0.8))
    blo = np.maximum(lo[:, None], np.minimum(ls, le) - 0.4).T.copy()  # [R][7]
    bhi = np.minimum(hi[:, None], np.maximum(ls, le) + 0.4).T.copy()
    rng = np.random.default_rng(4002)
    dev = torch.device("cuda:0")
    vox = torch.from_numpy(rng.uniform(-0.8, 0.8, (voxels, 3)).astype(np.float32)).to(dev)
    rg = list(pl.getRobotGeometryIds())
    cand = 256
    u = rng.uniform(0, 1, (regions, cand, 7)).astype(np.float32)
    samples = torch.from_numpy(blo[:, None, :] + u * (bhi - blo)[:, None, :]).to(dev).contiguous()
    flags = torch.empty(regions * cand, dtype=torch.uint8, device=dev)
    capi.check_collision_free_voxels(dp, samples, regions * cand, 7, 1, vox, voxels, 0.02, rg, flags)
    col = flags.view(regions, cand) == 0
    keep = col & (torch.cumsum(col.int(), dim=1) <= per)
    cols = samples[keep].contiguous()  # region-major
    seg = torch.nonzero(keep)[:, 0].to(torch.int32).contiguous()
    offsets = torch.zeros(regions + 1, dtype=torch.int32, device=dev)
    offsets[1:] = torch.cumsum(keep.sum(dim=1), dim=0).to(torch.int32)
    C = cols.shape[0]
    F0, cap = 14, 14 + per
    A0 = torch.zeros(regions, cap, 7, device=dev)
    b0 = torch.zeros(regions, cap, device=dev)
    for i in range(7):  # MakeBox rows, interleaved (-e_i, +e_i)
        A0[:, 2 * i, i], A0[:, 2 * i + 1, i] = -1.0, 1.0
        b0[:, 2 * i] = -torch.from_numpy(blo[:, i]).to(dev)
        b0[:, 2 * i + 1] = torch.from_numpy(bhi[:, i]).to(dev)
    nf0 = torch.full((regions,), F0, dtype=torch.int32, device=dev)
    A, b, nf = A0.clone(), b0.clone(), nf0.clone()
    dls = torch.from_numpy(np.ascontiguousarray(ls.T)).to(dev)
    dle = torch.from_numpy(np.ascontiguousarray(le.T)).to(dev)
    scol = torch.empty(C * 7, device=dev)
    proj = torch.empty(C * 7, device=dev)
    status = torch.empty(regions, dtype=torch.int32, device=dev)
    s = torch.cuda.current_stream()

    def run():
        A.copy_(A0)
        b.copy_(b0)
        nf.copy_(nf0)
        capi.edit_regions(dp, cols, seg, offsets, C, dls, dle, regions, A, b, nf, cap, 0.01, steps,
                          scol, proj, status, voxels=vox, num_voxels=voxels, radius=0.02,
                          robot_geoms=rg, stream=s)

    for _ in range(2):
        run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        run()
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    st = status.cpu().numpy()
    if (st < 0).any():
        raise RuntimeError(f"region editing failed for {(st < 0).sum()} regions")
    return {"regions": regions, "collisions": C, "bisection_steps": steps, "voxels": voxels,
            "ms": round(ms, 4), "regions_per_s": round(regions / (ms * 1e-3)),
            "collisions_per_s": round(C / (ms * 1e-3)),
            "checks_per_s": round(C * steps / (ms * 1e-3)),
            "faces_added": int(st.sum()), "faces_max": int(st.max())}


def distance_field_benchmark(grid=512, sites=20000, reps=10):
    """Distance field of a grid^3 voxel grid (voxel 1 cm) with `sites`
    occupied voxels at uniform random positions: one computeEDT (row pass +
    two column passes) timed with HIP events. Per voxel the passes read the
    uint8 occupancy once and read and write one int32 each (21 B), plus the
    column passes' stack reads."""
    import torch
    from csdecomp_amd import capi
    rng = np.random.default_rng(5001)
    pts = torch.from_numpy(rng.uniform(0, grid * 0.01, (sites, 3)).astype(np.float32)).cuda()
    f = capi.DistanceField([0.0, 0.0, 0.0], grid, 0.01)
    s = torch.cuda.current_stream()
    f.set_occupancy(pts, sites, s)
    f.compute(s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        f.compute(s)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    voxels = grid ** 3
    return {"grid": grid, "sites": sites, "ms": round(ms, 3),
            "voxels_per_s": round(voxels / (ms * 1e-3)),
            "pass_io_GBps": round(21 * voxels / (ms * 1e-3) / 1e9, 1)}


def voxel_maps_benchmark(nodes=20000, edges=2000, step=0.02):
    """DRM offline maps on the 7-DOF arm (48 boxes + 16 spheres, 32 robot
    spheres) over a 64^3 grid of 3.125 cm voxels (2 m cube, resolution
    1/64 m): the node map of `nodes` uniform configurations and the edge map
    of `edges` edges of length U[0.1, 0.4] rad at `step`, each one
    csd_voxel_collision_map call (count pass, host scan, fill pass)."""
    import time
    import torch
    from csdecomp_amd import capi, scenes
    pl = scenes.arm_scene(1001, 48, 16)
    dp = capi.DevicePlant(pl.getMinimalPlant())
    rg = list(pl.getSceneInspector().robot_geometry_ids)
    lo = np.asarray(pl.getPositionLowerLimits(), np.float32)
    hi = np.asarray(pl.getPositionUpperLimits(), np.float32)
    rng = np.random.default_rng(6001)
    q = rng.uniform(lo, hi, (nodes, 7)).astype(np.float32)
    d = rng.normal(size=(edges, 7))
    d *= rng.uniform(0.1, 0.4, (edges, 1)) / np.linalg.norm(d, axis=1, keepdims=True)
    q2 = np.clip(q[:edges] + d, lo, hi).astype(np.float32)
    allq = torch.from_numpy(np.concatenate([q, q2])).cuda()
    e = torch.from_numpy(np.stack([np.arange(edges), nodes + np.arange(edges)], 1)
                         .astype(np.int32)).cuda()
    center, size, res = [0.0, 0.0, 0.5], [2.0, 2.0, 2.0], 1.0 / 64
    V = 64 ** 3
    off = torch.empty(V + 1, dtype=torch.int64, device="cuda")
    res_out = {"grid": [64, 64, 64], "nodes": nodes, "edges": edges, "edge_step": step}
    for name, ed, ne in (("node_map", None, 0), ("edge_map", e, edges)):
        args = (dp, allq, nodes + edges, ed, ne, step, rg, center, size, res, off)
        total = capi.voxel_collision_map(*args)
        items = torch.empty(max(total, 1), dtype=torch.int32, device="cuda")
        capi.voxel_collision_map(*args, items=items, capacity=total)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        capi.voxel_collision_map(*args, items=items, capacity=total)
        dt = time.perf_counter() - t0
        res_out[name] = {"entries": total, "s": round(dt, 4),
                         "items_per_s": round((nodes if ed is None else edges) / dt)}
    return res_out


def eizo_benchmarks(segments=2048, cpu=True):
    """EI-ZO legs: single-edge inflation (7-DOF arm, 48 boxes + 16 spheres, 256
    segments, 4096 particles) and batched inflation of uneven sizes (14-DOF
    dual arm, 96 obstacles, 2048 segments, lengths log-U[0.05, 2] rad, budgets
    log-U[512, 16384]); plain rounds vs surrogate-guided rounds. Each
    surrogate is trained for its workload (2M labels, half around the
    segments, 6,000 Adam steps of 8,192) and its labelling and training
    time is counted in the guided regions/s."""
    from csdecomp_amd import scenes, surrogate
    res = {}
    arm = scenes.arm_scene(1001, 48, 16)
    s1, e1 = scenes.free_segments(arm, 256, 3001, (0.2, 1.5))
    m7, tr7 = surrogate.train_surrogate(arm, s1, e1)
    modes = ("plain", "plain_graph", "guided", "guided_graph", "guided_untrained", "guided_fallback",
             "guided_pruned", "guided_pruned_graph", "guided_pruned_untrained")
    res["single_edge_arm7"] = eizo_leg("single", arm, s1, e1, [], 4096, modes, m7,
                                       train_s=tr7["label_s"] + tr7["train_s"])
    res["single_edge_arm7"]["surrogate"] = tr7
    if cpu:
        res["single_edge_arm7"]["cpu_baseline"] = eizo_cpu_baseline(arm, s1, e1)
    dual = scenes.dual_arm_scene(1002, 72, 24)
    s2, e2 = scenes.free_segments(dual, segments, 3002, (0.05, 2.0), log_uniform=True)
    b2 = scenes.log_uniform_budgets(3003, segments)
    m14, tr14 = surrogate.train_surrogate(dual, s2, e2)
    t14 = tr14["label_s"] + tr14["train_s"]
    res["batched_dualarm14"] = eizo_leg("batched", dual, s2, e2, b2, 16384, modes, m14,
                                        train_s=t14)
    res["batched_dualarm14"]["surrogate"] = tr14
    # room for 1,000 faces, where plain certifies the majority of the
    # regions: the first 128 segments (a surrogate trained on them)
    k = min(128, segments)
    sk, ek = np.asfortranarray(s2[:, :k]), np.asfortranarray(e2[:, :k])
    mk, trk = surrogate.train_surrogate(dual, sk, ek)
    res["batched_dualarm14_1000faces"] = eizo_leg(
        "batched", dual, sk, ek, b2[:k], 16384, ("plain", "guided", "guided_untrained"), mk,
        face_budget=1000, train_s=trk["label_s"] + trk["train_s"])
    res["batched_dualarm14_1000faces"]["surrogate"] = trk
    return res


def eizo_plain_benchmarks(segments=256):
    """Plain EI-ZO only (no surrogate, no training), `segments` segments of
    each workload: the leg counter passes run over."""
    from csdecomp_amd import scenes
    arm = scenes.arm_scene(1001, 48, 16)
    s1, e1 = scenes.free_segments(arm, segments, 3001, (0.2, 1.5))
    dual = scenes.dual_arm_scene(1002, 72, 24)
    s2, e2 = scenes.free_segments(dual, segments, 3002, (0.05, 2.0), log_uniform=True)
    b2 = scenes.log_uniform_budgets(3003, segments)
    return {"single_edge_arm7": eizo_leg("single", arm, s1, e1, [], 4096, ("plain",), None),
            "batched_dualarm14": eizo_leg("batched", dual, s2, e2, b2, 16384, ("plain",), None)}


def graph_replay_benchmark(dp, dq, n, ref_flags, launches=50, replays=5):
    """The headline check captured as one hipGraph of `launches` launches
    (torch.cuda.graph over the C-ABI call, which only launches kernels) and
    replayed `replays` times, the output cleared eagerly between replays;
    every replay's flags must equal the eager flags."""
    import torch
    from csdecomp_amd import capi
    out = torch.empty(n, dtype=torch.uint8, device=dq.device)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        capi.check_collision_free(dp, dq, n, 1, n, out, s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(launches):
            capi.check_collision_free(dp, dq, n, 1, n, out, s)
    torch.cuda.synchronize()
    times, identical = [], True
    for _ in range(replays):
        out.zero_()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        g.replay()
        b.record()
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b))
        identical = identical and bool(torch.equal(out, ref_flags))
    ms = float(np.median(times)) / launches
    del g
    return {"launches_per_graph": launches, "replays": replays, "ms_per_launch": ms,
            "configs_per_s": n / (ms * 1e-3), "flags_identical_to_eager": identical}


def large_batch_benchmark(plant, dp, stream, sizes=(262144, 1048576, 4194304, 16777216), steps=20):
    """The headline check on larger batches of the same workload (configs
    seed 2000), each autotuned first: from a few hundred thousand
    configurations on, the tuner picks the screened checker."""
    import torch
    from csdecomp_amd import capi, scenes
    res = {}
    for n in sizes:
        q = scenes.uniform_configs(2000, n, plant.getPositionLowerLimits(),
                                   plant.getPositionUpperLimits())
        dq = torch.from_numpy(np.ascontiguousarray(q)).to("cuda:0")
        out = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        shape = capi.tune_check(dp, dq, n, 1, n, out, stream)
        for _ in range(3):
            capi.check_collision_free(dp, dq, n, 1, n, out, stream)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for _ in range(steps):
            capi.check_collision_free(dp, dq, n, 1, n, out, stream)
        t1.record(stream)
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / steps
        res[str(n)] = {"configs_per_s": n / (ms * 1e-3), "ms_per_step": ms,
                       "checker_shape": shape}
        del dq, out
    return res


def _large_batches_leg():
    import torch
    from csdecomp_amd import capi, scenes
    plant = scenes.arm_scene(1000, 16, 0)
    dp = capi.DevicePlant(plant.getMinimalPlant())
    return large_batch_benchmark(plant, dp, torch.cuda.current_stream())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--num-configs", type=int, default=65536)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-eizo", action="store_true", help="skip the EI-ZO regions/s legs")
    ap.add_argument("--no-large", action="store_true", help="skip the large-batch checker leg")
    ap.add_argument("--eizo-segments", type=int, default=2048)
    ap.add_argument("--shape", type=int, default=0,
                    help="checker shape for the headline (0: autotune; 1, 2, 4, 11, 12, 14, 20: "
                         "that shape, e.g. to profile it)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    ap.add_argument("--leg", choices=["eizo", "eizo_plain", "mlp", "roadmap", "edit_regions", "distance_field",
                                      "voxel_maps", "large_batches"],
                    help="only run this secondary leg and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.leg:
        from csdecomp_amd import capi
        capi.init(0)
        fn = {"eizo": lambda: eizo_benchmarks(args.eizo_segments, not args.no_cpu),
              "eizo_plain": lambda: eizo_plain_benchmarks(args.eizo_segments),
              "mlp": mlp_benchmark,
              "roadmap": roadmap_benchmark, "edit_regions": edit_regions_benchmark,
              "distance_field": distance_field_benchmark,
              "voxel_maps": voxel_maps_benchmark,
              "large_batches": _large_batches_leg}[args.leg]
        print(json.dumps({args.leg: fn()}), flush=True)
        return
    if args.copy_calibration:
        import torch
        src = torch.ones(args.copy_calibration * (1 << 18), device="cuda:0")  # 4-byte floats
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the checker shards "
              "configurations across GPUs with no exchange, which this build does not launch.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from csdecomp_amd import capi, scenes

    capi.init(0)
    dev = torch.device("cuda:0")
    plant = scenes.arm_scene(1000, 16, 0)  # scene seed 1000
    mp = plant.getMinimalPlant()
    n = args.num_configs
    q = scenes.uniform_configs(2000, n, plant.getPositionLowerLimits(), plant.getPositionUpperLimits())
    dq = torch.from_numpy(np.ascontiguousarray(q)).to(dev)  # [7][N] structure of arrays
    out = torch.empty(n, dtype=torch.uint8, device=dev)
    dp = capi.DevicePlant(mp)
    stream = torch.cuda.current_stream()

    def step():
        capi.check_collision_free(dp, dq, n, 1, n, out, stream)

    # autotuning (outside the timed region): waves per group of 64
    # configurations, one or two passes, or the screened checker
    split = capi.tune_check(dp, dq, n, 1, n, out, stream, force=args.shape)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(args.steps):
        step()
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps

    # one launch per step, back to back on one stream: the HIP events around
    # the timed region give the kernel's average launch duration
    kern_ms = ms
    # a launch on its own (events around each launch) for reference
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(min(args.steps, 50))]
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    single_ms = float(np.median([a.elapsed_time(b) for a, b in evs]))

    free_frac = float(out.float().mean().item())
    f_cfg = algorithmic_flops_per_config(mp, plant)
    achieved_tflops = n * f_cfg / (kern_ms * 1e-3) / 1e12
    # HBM bytes per launch of the shape timed here, from its rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (null for a shape not profiled)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_check_arm7_16box_traffic.json")
    if os.path.exists(tfile) and n == 65536:
        with open(tfile) as f:
            traffic = json.load(f).get("shapes", {}).get(str(split), {}).get("hbm_bytes_per_launch")

    line = {
        "metric": "collision-checked configs/sec",
        "value": n / (ms * 1e-3),
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene 1000, configs seed 2000)",
        "config": {"workload": "FK + collision check, 7-DOF arm, 32 spheres x 16 boxes, 784 pairs",
                   "num_configs": n, "num_collision_pairs": mp.num_collision_pairs,
                   "layout": "SoA [7][N] f32 in HBM", "collision_free_fraction": round(free_frac, 4),
                   "kernel": ("k_check<fast 8x4>" if split < 10 else
                              "k_check2<fast 8x4> (self pairs, then obstacles for survivors)"
                              if split < 20 else
                              "k_screen + k_finish <fast 8x4> (screened pairs, then the full "
                              "check of the undecided)"),
                   "checker_shape": split,
                   "waves_per_64_configs": split if split < 10 else split - 10},
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": FP32_VECTOR_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / FP32_VECTOR_PEAK_TFLOPS,
                     "traffic": traffic, "flops_per_config": f_cfg, "kernel_ms": kern_ms,
                     "single_launch_ms": single_ms,
                     "algorithmic_bytes": n * (4 * plant_dof(plant) + 1)},
    }
    line["graph_replay"] = graph_replay_benchmark(dp, dq, n, out.clone())
    if not args.no_large:
        line["large_batches"] = large_batch_benchmark(plant, dp, stream)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mp, q, args.cpu_seconds)
    if not args.no_eizo:
        line["eizo"] = eizo_benchmarks(args.eizo_segments, not args.no_cpu)
        line["surrogate_mlp"] = mlp_benchmark()
        line["roadmap"] = roadmap_benchmark()
        line["edit_regions"] = edit_regions_benchmark()
        line["distance_field"] = distance_field_benchmark()
        line["voxel_maps"] = voxel_maps_benchmark()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
