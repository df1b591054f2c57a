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


def cpu_baseline(mp, q, seconds):
    """CPU restatement of checkCollisionFreeVoxels with no voxels (early exit,
    silent), OpenMP over configurations on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    op = oracle.Plant(mp)
    threads = os.cpu_count() or 1
    n = q.shape[1]
    oracle.check(op, q[:, :1024], oracle.RULE_CPU, threads)  # warm up
    done, t0 = 0, time.perf_counter()
    while True:
        oracle.check(op, q, oracle.RULE_CPU, threads)
        done += n
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": done / el, "unit": "configs/s", "cores": threads, "kind": "port",
            "sample": f"{done} configs ({done // n} passes over the {n}-config batch), {el:.1f} s, "
                      f"CPU rule of collision_checker.cpp with early exit"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--num-configs", type=int, default=65536)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
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

    # per-launch kernel duration (events around each launch, same stream)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(min(args.steps, 50))]
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))

    free_frac = float(out.float().mean().item())
    f_cfg = algorithmic_flops_per_config(mp, plant)
    achieved_tflops = n * f_cfg / (kern_ms * 1e-3) / 1e12
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_check_arm7_16box_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

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
                   "kernel": "k_check<fast 8x4>"},
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": FP32_VECTOR_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / FP32_VECTOR_PEAK_TFLOPS,
                     "traffic": traffic, "flops_per_config": f_cfg, "kernel_ms": kern_ms},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mp, q, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
