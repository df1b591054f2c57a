"""CPU restatement of the signed-distance surrogate (csd_mlp_* of the C-ABI).

Test infrastructure only: tests/ use it as the checker of the GPU network.
Nothing in csdecomp_amd/ imports it.

The network (no counterpart in the reference):
    x   = 2 (q - lower) / (upper - lower) - 1, zero-padded to 32 features
    a_l = r(SiLU(W_l a_{l-1} + b_l)),  l = 1..4, width 256
    y   = r(w5) . a_4 + b_5
and the mean-squared-error loss over a batch. One implementation serves two
purposes through the rounding function r applied to weights, activations and
back-propagated deltas:
  * r = identity: the exact float64 network; its backward pass is pinned by
    central finite differences (tests/test_oracle_mlp.py);
  * r = bf16: the device's rounding points (bf16 weights, activations and
    deltas; FP32 biases and bias gradients), float64 in between, so GPU
    results match up to FP32 accumulation-order noise.
"""
import numpy as np

WIDTH, IN_PAD, HIDDEN = 256, 32, 4


def _offsets():
    off, o = {}, 0
    for l in range(1, HIDDEN + 1):
        k = IN_PAD if l == 1 else WIDTH
        off[f"W{l}"] = (o, (WIDTH, k))
        o += WIDTH * k
        off[f"b{l}"] = (o, (WIDTH,))
        o += WIDTH
    off["w5"] = (o, (WIDTH,))
    o += WIDTH
    off["b5"] = (o, (1,))
    o += 1
    return off, o


OFFSETS, NUM_PARAMS = _offsets()


def unpack(p):
    p = np.asarray(p)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFFSETS.items()}


def pack(d):
    p = np.zeros(NUM_PARAMS, np.float64)
    for k, (o, s) in OFFSETS.items():
        p[o:o + int(np.prod(s))] = np.asarray(d[k], np.float64).reshape(-1)
    return p


def bf16(x):
    """Round to bfloat16, nearest-even (values returned as float64)."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, u | 0x00400000, r) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def identity(x):
    return np.asarray(x, np.float64)


def normalise(q, lower, upper):
    """x = (q - lower) * scale - 1 with scale = 2 / (upper - lower), the
    device's FP32 expression (product and subtraction fused), zero-padded to
    32 features. q: [B][in_dim]."""
    q = np.asarray(q, np.float32)
    lo = np.asarray(lower, np.float32)
    sc = (np.float32(2.0) / (np.asarray(upper, np.float32) - lo)).astype(np.float32)
    d = (q - lo).astype(np.float32)
    x = (d.astype(np.float64) * sc.astype(np.float64) - 1.0).astype(np.float32)
    out = np.zeros((q.shape[0], IN_PAD), np.float64)
    out[:, :q.shape[1]] = x
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, r=bf16):
    """Returns (y [B], cache). x: normalised inputs [B][32]."""
    P = unpack(np.asarray(p, np.float64))
    a = [r(x)]
    z = [None]
    for l in range(1, HIDDEN + 1):
        zl = a[-1] @ r(P[f"W{l}"]).T + P[f"b{l}"]
        z.append(zl)
        a.append(r(_silu(zl)))
    y = a[-1] @ r(P["w5"]) + P["b5"][0]
    return y, (P, a, z)


def loss_and_grads(p, x, target, r=bf16):
    """Mean squared error and its gradient (flat, parameter layout)."""
    y, (P, a, z) = forward(p, x, r)
    B = x.shape[0]
    t = np.asarray(target, np.float64)
    loss = np.mean((y - t) ** 2)
    dy = 2.0 * (y - t) / B
    g = {"w5": dy @ a[HIDDEN], "b5": np.array([dy.sum()])}
    d = np.outer(dy, r(P["w5"])) * _dsilu(z[HIDDEN])
    for l in range(HIDDEN, 0, -1):
        g[f"b{l}"] = d.sum(axis=0)
        dr = r(d)
        g[f"W{l}"] = dr.T @ a[l - 1]
        if l > 1:
            d = (dr @ r(P[f"W{l}"])) * _dsilu(z[l - 1])
    return loss, pack(g)
