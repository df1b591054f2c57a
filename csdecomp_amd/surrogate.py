"""Training of the signed-distance surrogate that guides EI-ZO
(CudaEdgeInflator.set_guidance).

The network (csd_mlp: in_dim -> 256 -> 256 -> 256 -> 256 -> 1, SiLU, BF16
weights and activations, FP32 accumulation) learns the signed distance of
the plant's robot geometry to the obstacles, clipped to +-0.2 m and scaled
to +-1, positive = free. Labels come from the exact checker on the GPU
(csd_signed_distance); training runs the library's own MFMA forward and
backward kernels with Adam (csd_mlp_train_step). The reference has no
counterpart: its checker returns booleans only.

EI-ZO queries the network inside polytopes grown around the segments, not
uniformly over the joint box. In a cluttered scene most of the joint box
collides, and a network fitted to uniform labels alone predicts a near
constant "in collision" there. Labels are therefore drawn half uniformly
over the joint limits and half around the segments: a uniform point of a
uniformly chosen segment plus isotropic Gaussian noise of `sigma` rad,
clipped to the limits.
"""
import time

import numpy as np


def sample_configurations(line_starts, line_ends, lower, upper, n, near_fraction=0.5,
                          sigma=0.25, seed=7, device="cuda"):
    """n configurations [n][dim] float32 on `device`: round(near_fraction * n)
    around the segments (columns of line_starts / line_ends, dim x R), the
    rest uniform over [lower, upper], shuffled. Seeded torch generator on
    the device."""
    import torch
    lo = torch.as_tensor(np.asarray(lower, np.float32), device=device)
    hi = torch.as_tensor(np.asarray(upper, np.float32), device=device)
    dim = lo.numel()
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    n_near = 0
    if line_starts is not None and near_fraction > 0:
        ls = torch.as_tensor(np.asarray(line_starts, np.float32), device=device).t().contiguous()
        le = torch.as_tensor(np.asarray(line_ends, np.float32), device=device).t().contiguous()
        if ls.shape[1] != dim or le.shape != ls.shape:
            raise ValueError("line_starts / line_ends must be dim x R")
        n_near = int(round(near_fraction * n))
    q = torch.empty(n, dim, device=device)
    u = torch.rand(n - n_near, dim, device=device, generator=g)
    q[n_near:] = lo + u * (hi - lo)
    if n_near:
        k = torch.randint(0, ls.shape[0], (n_near,), device=device, generator=g)
        t = torch.rand(n_near, 1, device=device, generator=g)
        p = ls[k] + t * (le[k] - ls[k])
        p += sigma * torch.randn(n_near, dim, device=device, generator=g)
        q[:n_near] = torch.minimum(torch.maximum(p, lo), hi)
    perm = torch.randperm(n, device=device, generator=g)
    return q[perm].contiguous()


def train_surrogate(plant, line_starts=None, line_ends=None, n=2_000_000, steps=6000,
                    batch=8192, lr=1e-3, near_fraction=0.5, sigma=0.25, seed=7,
                    device="cuda"):
    """Labels n configurations of `plant` (a csdecomp_amd Plant) with the
    exact signed distance and trains a fresh network for `steps` Adam steps
    of `batch` rows (consecutive slices of the labelled set, wrapping).
    Returns (capi.Mlp, stats) with stats: labels, label_s, train_steps,
    train_s, final_mse (the last batch's loss)."""
    import torch
    from . import capi
    if batch % 128 or batch > n:
        raise ValueError("batch must be a multiple of 128 and at most n")
    mp = plant.getMinimalPlant()
    lo = np.asarray(plant.getPositionLowerLimits(), np.float32)
    hi = np.asarray(plant.getPositionUpperLimits(), np.float32)
    dof = len(lo)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    q = sample_configurations(line_starts, line_ends, lo, hi, n, near_fraction, sigma, seed,
                              device)
    sd = torch.empty(n, device=device)
    dplant = capi.DevicePlant(mp)  # alive until the labels are computed
    capi.signed_distance(dplant, q, n, dof, 1, sd)
    target = (sd.clamp(-0.2, 0.2) / 0.2).contiguous()
    torch.cuda.synchronize(device)
    t1 = time.perf_counter()
    model = capi.Mlp(dof, lo, hi, seed)
    loss = torch.empty(1, device=device)
    for s in range(steps):
        k = (s * batch) % (n - batch + 1)
        model.train_step(q[k:k + batch], target[k:k + batch], batch, lr, loss)
    torch.cuda.synchronize(device)
    t2 = time.perf_counter()
    del dplant
    return model, {"labels": n, "label_s": round(t1 - t0, 3), "train_steps": steps,
                   "train_s": round(t2 - t1, 3), "final_mse": round(loss.item(), 5)}
