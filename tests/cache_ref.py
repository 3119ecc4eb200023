"""NumPy model of the radiance cache's BF16 arithmetic (test infrastructure).

The GPU kernels round the encoded input, the weights, every hidden activation
and every back-propagated pre-activation gradient to bfloat16 (round to nearest
even) and accumulate in FP32.  `forward_bf16` / `loss_grad_bf16` repeat those
roundings with FP64 accumulation, so they differ from the GPU only by
accumulation order and the rare bfloat16 rounding it flips; with `bf16=False`
they are plain FP64 and must equal the oracle (pinned in
test_cache_reference.py).  Parameter layout: per layer W (out x in, row-major)
then b, layers 81->64->64->64->64->3.
"""
import numpy as np

FAN_IN = (81, 64, 64, 64, 64)
FAN_OUT = (64, 64, 64, 64, 3)
OFF_W = (0, 5248, 9408, 13568, 17728)
OFF_B = (5184, 9344, 13504, 17664, 17920)
N_PARAMS = 17923


def bf16(x) -> np.ndarray:
    """Round FP32 values to bfloat16 (nearest, ties to even), returned as FP32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def weights(p, l):
    return np.asarray(p[OFF_W[l]:OFF_W[l] + FAN_IN[l] * FAN_OUT[l]]).reshape(FAN_OUT[l], FAN_IN[l])


def bias(p, l):
    return np.asarray(p[OFF_B[l]:OFF_B[l] + FAN_OUT[l]])


def encode(q, lo, hi) -> np.ndarray:
    """81 inputs of each 12-float query: sin/cos(2^f pi p_a) for f < 12 with
    p = (pos - lo) * (1 / (hi - lo)) in FP32 as the kernels form it, then
    direction, normal, albedo."""
    q = np.asarray(q, np.float32).reshape(-1, 12)
    lo32 = np.asarray(lo, np.float32)
    inv = np.float32(1.0) / (np.asarray(hi, np.float32) - lo32)
    p01 = ((q[:, :3] - lo32) * inv).astype(np.float32).astype(np.float64)
    enc = np.zeros((q.shape[0], 81))
    for a in range(3):
        for f in range(12):
            arg = np.pi * (2.0 ** f) * p01[:, a]
            enc[:, 24 * a + 2 * f] = np.sin(arg)
            enc[:, 24 * a + 2 * f + 1] = np.cos(arg)
    enc[:, 72:81] = q[:, 3:12]
    return enc


def _round(x, on):
    return bf16(x.astype(np.float32)).astype(np.float64) if on else x


def _forward(p, enc, on):
    acts, pre = [_round(enc, on)], []
    for l in range(5):
        z = acts[-1] @ _round(weights(p, l).astype(np.float64), on).T + bias(p, l).astype(np.float64)
        pre.append(z)
        if l < 4:
            acts.append(_round(np.maximum(z, 0.0), on))
    return acts, pre


def forward_bf16(p, enc, bf16_on=True) -> np.ndarray:
    """Output pre-activations z (radiance = exp z), shape (n, 3)."""
    return _forward(p, enc, bf16_on)[1][4]


def loss_grad_bf16(p, enc, targets, eps=1e-2, bf16_on=True):
    """Relative L2 loss (y - t)^2 / (sg(y)^2 + eps) averaged over samples and
    channels, and its gradient (FP64 array of N_PARAMS)."""
    acts, pre = _forward(p, enc, bf16_on)
    n = enc.shape[0]
    y = np.exp(pre[4])
    t = np.asarray(targets, np.float64).reshape(n, 3)
    r, den, norm = y - t, y * y + eps, 1.0 / (n * 3)
    loss = float((r * r / den * norm).sum())
    d = _round(2.0 * r / den * norm * y, bf16_on)
    g = np.zeros(N_PARAMS)
    for l in range(4, -1, -1):
        g[OFF_W[l]:OFF_W[l] + FAN_IN[l] * FAN_OUT[l]] = (d.T @ acts[l]).reshape(-1)
        g[OFF_B[l]:OFF_B[l] + FAN_OUT[l]] = d.sum(0)
        if l > 0:
            dn = d @ _round(weights(p, l).astype(np.float64), bf16_on)
            d = _round(np.where(pre[l - 1] > 0.0, dn, 0.0), bf16_on)
    return loss, g


def random_params(seed, out_scale=0.5) -> np.ndarray:
    """He-uniform weights (output layer scaled by out_scale), small biases."""
    r = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for l in range(5):
        lim = np.sqrt(6.0 / FAN_IN[l]) * (out_scale if l == 4 else 1.0)
        p[OFF_W[l]:OFF_W[l] + FAN_IN[l] * FAN_OUT[l]] = r.uniform(-lim, lim, FAN_IN[l] * FAN_OUT[l])
        p[OFF_B[l]:OFF_B[l] + FAN_OUT[l]] = r.uniform(-0.1, 0.1, FAN_OUT[l])
    return p


def random_queries(n, lo, hi, seed) -> np.ndarray:
    """Positions uniform in the box, unit directions and normals, albedo in [0, 1)."""
    r = np.random.default_rng(seed)
    pos = r.uniform(lo, hi, (n, 3))
    d = r.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nn = r.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    return np.concatenate([pos, d, nn, r.uniform(0.0, 1.0, (n, 3))], 1).astype(np.float32)


def smooth_target(q, lo, hi) -> np.ndarray:
    """A smooth positive radiance-like function of position, direction and albedo."""
    q = np.asarray(q, np.float64)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    p01 = (q[:, :3] - lo) / (hi - lo)
    base = 0.3 + 0.25 * np.sin(2 * np.pi * p01[:, 0]) * np.cos(np.pi * p01[:, 2]) + 0.2 * p01[:, 1]
    return (base[:, None] * (0.2 + q[:, 9:12]) * (1.0 + 0.3 * np.clip(q[:, 4:5], -1, 1))).astype(np.float32)
