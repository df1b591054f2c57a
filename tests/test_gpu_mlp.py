"""GPU parity of the signed-distance surrogate: exact labels against the FP64
oracle, MFMA inference and gradients against the CPU restatement with the
device's bf16 rounding points, and a training run that must learn."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp as omlp  # noqa: E402


def _scene(name):
    from csdecomp_amd import scenes
    return {"arm7_16box": lambda: scenes.arm_scene(1000, 16, 0),
            "arm7_48box16sph": lambda: scenes.arm_scene(1001, 48, 16),
            "dualarm14": lambda: scenes.dual_arm_scene(1002, 72, 24)}[name]()


@pytest.mark.parametrize("name", ["arm7_16box", "arm7_48box16sph", "dualarm14"])
def test_signed_distance_vs_oracle(gpu, orc, name):
    import torch
    from csdecomp_amd import capi, scenes
    pl = _scene(name)
    mp = pl.getMinimalPlant()
    n = 4000
    q = scenes.uniform_configs(77, n, pl.getPositionLowerLimits(), pl.getPositionUpperLimits())
    dp = capi.DevicePlant(mp)
    dq = torch.from_numpy(np.ascontiguousarray(q)).to(gpu)  # [dof][n]
    out = torch.empty(n, device=gpu)
    capi.signed_distance(dp, dq, n, 1, n, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    ref = orc.min_sd(orc.Plant(mp), q)
    # FP32 arithmetic on a metre scale
    assert np.max(np.abs(got - ref)) < 2e-5, np.max(np.abs(got - ref))
    # sign agrees with the collision flags away from contact
    flags = orc.check(orc.Plant(mp), q, orc.RULE_GPU)
    far = np.abs(ref) > 1e-5
    assert np.all((got[far] >= 0) == (flags[far] == 1))


def test_signed_distance_rejects_non_sphere_plants(gpu, csd):
    import torch
    from csdecomp_amd import capi
    from conftest import parse
    mp = parse("3dofrobot.urdf").get_minimal_plant()  # moving boxes
    dp = capi.DevicePlant(mp)
    q = torch.zeros(3, device=gpu)
    out = torch.empty(1, device=gpu)
    with pytest.raises(capi.CsdError) as e:
        capi.signed_distance(dp, q, 1, 3, 1, out)
    assert e.value.status == 5  # CSD_ERR_UNSUPPORTED


def _model(in_dim, seed):
    from csdecomp_amd import capi
    lo, hi = np.full(in_dim, -np.pi, np.float32), np.full(in_dim, np.pi, np.float32)
    return capi.Mlp(in_dim, lo, hi, seed), lo, hi


@pytest.mark.parametrize("in_dim", [2, 7, 14])
def test_predict_matches_oracle(gpu, in_dim):
    import torch
    m, lo, hi = _model(in_dim, 11)
    p = m.get_params().astype(np.float64)
    # the padded input columns start at zero
    assert np.all(omlp.unpack(p)["W1"][:, in_dim:] == 0)
    rng = np.random.default_rng(in_dim)
    n = 1000  # not a multiple of the 128-row tile
    q = rng.uniform(-np.pi, np.pi, (n, in_dim)).astype(np.float32)
    dq = torch.from_numpy(q).to(gpu)
    out = torch.empty(n, device=gpu)
    m.predict(dq, n, in_dim, 1, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref, _ = omlp.forward(p, omlp.normalise(q, lo, hi))
    err = np.abs(got - ref)
    scale = 1.0 + np.abs(ref).max()
    assert err.max() < 2e-2 * scale and err.mean() < 2e-3 * scale, (err.max(), err.mean())
    # larger output values after perturbing the head: still consistent
    p2 = p.copy()
    o, _ = omlp.OFFSETS["w5"]
    p2[o:o + 256] *= 8.0
    m.set_params(p2.astype(np.float32))
    m.predict(dq, n, in_dim, 1, out)
    torch.cuda.synchronize()
    ref2, _ = omlp.forward(p2.astype(np.float32).astype(np.float64), omlp.normalise(q, lo, hi))
    err2 = np.abs(out.cpu().numpy() - ref2)
    assert err2.max() < 2e-2 * (1.0 + np.abs(ref2).max())


def test_gradients_match_oracle(gpu):
    import torch
    in_dim, B = 7, 512
    m, lo, hi = _model(in_dim, 5)
    rng = np.random.default_rng(8)
    q = rng.uniform(-np.pi, np.pi, (B, in_dim)).astype(np.float32)
    t = rng.uniform(-1, 1, B).astype(np.float32)
    grads = torch.empty(m.num_params, device=gpu)
    loss = torch.empty(1, device=gpu)
    m.gradients(torch.from_numpy(q).to(gpu), torch.from_numpy(t).to(gpu), B, grads, loss)
    torch.cuda.synchronize()
    g = grads.cpu().numpy().astype(np.float64)
    p = m.get_params().astype(np.float64)
    ref_loss, ref = omlp.loss_and_grads(p, omlp.normalise(q, lo, hi), t)
    assert abs(loss.item() - ref_loss) < 1e-3 * (1 + ref_loss)
    for name, (o, s) in omlp.OFFSETS.items():
        k = int(np.prod(s))
        a, b = g[o:o + k], ref[o:o + k]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 3e-2, (name, rel)
    # a parameter step against the gradient lowers the exact loss
    p2 = p - 0.05 * g / np.linalg.norm(g)
    l2, _ = omlp.loss_and_grads(p2, omlp.normalise(q, lo, hi), t, r=omlp.identity)
    l1, _ = omlp.loss_and_grads(p, omlp.normalise(q, lo, hi), t, r=omlp.identity)
    assert l2 < l1


def test_gradients_across_batch_sizes(gpu):
    """Gradient calls in a row with batches 384, 128 (the workspace is carved
    again for a smaller batch), 1024 (it grows) and 384 again, plus an Adam
    step in between: each call's gradient and loss match the oracle's for its
    own batch, so the gradient accumulators start every step at zero."""
    import torch
    in_dim = 7
    m, lo, hi = _model(in_dim, 9)
    rng = np.random.default_rng(21)
    for i, B in enumerate((384, 128, 1024, 384)):
        q = rng.uniform(-np.pi, np.pi, (B, in_dim)).astype(np.float32)
        t = rng.uniform(-1, 1, B).astype(np.float32)
        dq, dt = torch.from_numpy(q).to(gpu), torch.from_numpy(t).to(gpu)
        grads = torch.empty(m.num_params, device=gpu)
        loss = torch.empty(1, device=gpu)
        m.gradients(dq, dt, B, grads, loss)
        torch.cuda.synchronize()
        p = m.get_params().astype(np.float64)
        ref_loss, ref = omlp.loss_and_grads(p, omlp.normalise(q, lo, hi), t)
        assert abs(loss.item() - ref_loss) < 1e-3 * (1 + ref_loss), (B, loss.item(), ref_loss)
        g = grads.cpu().numpy().astype(np.float64)
        for name, (o, s) in omlp.OFFSETS.items():
            k = int(np.prod(s))
            a, b = g[o:o + k], ref[o:o + k]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
            assert rel < 3e-2, (B, name, rel)
        if i == 1:
            m.train_step(dq, dt, B, 1e-3)  # an Adam step between two gradient calls
            torch.cuda.synchronize()


def test_batch_argument_checks(gpu):
    import torch
    from csdecomp_amd import capi
    import ctypes
    m, _, _ = _model(7, 1)
    q = torch.zeros(256, 7, device=gpu)
    t = torch.zeros(256, device=gpu)
    with pytest.raises(ValueError):
        m.train_step(q[:100], t, 100, 1e-3)  # not a multiple of 128
    # the C-ABI refuses it too when called directly
    rc = capi.lib().csd_mlp_train_step(m.handle, ctypes.c_void_p(q.data_ptr()),
                                       ctypes.c_void_p(t.data_ptr()), 100, 1e-3, None, None)
    assert rc == 1  # CSD_ERR_INVALID_ARGUMENT
    with pytest.raises(TypeError):
        m.train_step(q.cpu(), t, 128, 1e-3)  # host memory
    with pytest.raises(TypeError):
        m.train_step(q.double(), t, 128, 1e-3)
    with pytest.raises(ValueError):
        m.train_step(q[:, :6], t, 128, 1e-3)  # in_dim columns
    with pytest.raises(ValueError):
        m.train_step(torch.zeros(7, 256, device=gpu).t(), t, 128, 1e-3)  # not contiguous
    with pytest.raises(ValueError):
        m.train_step(q, t[:64], 128, 1e-3)  # too few targets
    m.train_step(q, t, 256, 1e-3)
    torch.cuda.synchronize()


def test_surrogate_outlives_the_inflator_it_guides(gpu):
    """The inflator keeps the Mlp object alive, and the C-ABI refuses to
    destroy a model that an EI-ZO workspace holds."""
    import gc
    import weakref
    import torch
    from csdecomp_amd import capi, scenes
    import csdecomp_amd as csd
    pl = scenes.arm_scene(1001, 48, 16)
    lo = np.asarray(pl.getPositionLowerLimits(), np.float32)
    hi = np.asarray(pl.getPositionUpperLimits(), np.float32)
    dom = csd.HPolyhedron()
    dom.MakeBox(lo, hi)
    opts = csd.EizoOptions()
    opts.num_particles = 256
    opts.max_iterations = 3
    inf = csd.CudaEdgeInflator(pl.getMinimalPlant(), list(pl.getRobotGeometryIds()), opts, dom,
                               max_regions=2)
    m = capi.Mlp(7, lo, hi, 3)
    inf.set_guidance(pilot_samples=512, pilot_factor=10.0, surrogate=m, keep_particles=64)
    assert capi.lib().csd_mlp_destroy(m.handle) != 0  # attached: refused
    ref = weakref.ref(m)
    del m
    gc.collect()
    assert ref() is not None  # held by the inflator
    ls = np.zeros((7, 1), np.float32)
    le = np.full((7, 1), 0.05, np.float32)
    regions, infos = inf.inflateEdges(ls, le, None, 0.0, [256])
    assert infos[0].iterations >= 1
    inf.set_guidance()  # detached: the model may now be destroyed
    m = ref()
    assert capi.lib().csd_mlp_destroy(m.handle) == 0
    m.handle = None
    del m
    del inf
    gc.collect()
    assert ref() is None  # released with the inflator
    torch.cuda.synchronize()


def test_training_learns_signed_distance(gpu):
    """Labels from the exact checker on the 64-obstacle arm scene; Adam
    (lr 1e-3, batch 8192) must cut the loss and predict the collision sign."""
    import torch
    from csdecomp_amd import capi, scenes
    pl = _scene("arm7_48box16sph")
    mp = pl.getMinimalPlant()
    lo, hi = np.asarray(pl.getPositionLowerLimits(), np.float32), np.asarray(pl.getPositionUpperLimits(), np.float32)
    n, B = 262144, 8192
    rng = np.random.default_rng(21)
    q = rng.uniform(lo, hi, (n + 16384, 7)).astype(np.float32)
    dq = torch.from_numpy(q).to(gpu)
    dp = capi.DevicePlant(mp)
    sd = torch.empty(n + 16384, device=gpu)
    capi.signed_distance(dp, dq, n + 16384, 7, 1, sd)
    target = (sd.clamp(-0.2, 0.2) / 0.2).contiguous()
    m = capi.Mlp(7, lo, hi, 3)
    loss = torch.empty(1, device=gpu)
    losses = []
    for step in range(400):
        k = (step * B) % n
        m.train_step(dq[k:k + B], target[k:k + B], B, 1e-3, loss)
        if step % 50 == 0 or step == 399:
            torch.cuda.synchronize()
            losses.append(loss.item())
    assert losses[-1] < 0.5 * losses[0], losses
    pred = torch.empty(16384, device=gpu)
    m.predict(dq[n:], 16384, 7, 1, pred)
    torch.cuda.synchronize()
    p, t = pred.cpu().numpy(), target[n:].cpu().numpy()
    far = np.abs(t) > 0.25  # more than 5 cm from contact
    acc = np.mean((p[far] > 0) == (t[far] > 0))
    assert acc > 0.85, acc
