"""Pins for the NumPy BF16 model of the radiance cache (CPU only).

With the bfloat16 roundings switched off the model must equal the FP64 oracle;
with them on it must stay within the BF16 error the GPU tolerances allow, and
a planted layout mistake must land far outside those tolerances.
"""
import numpy as np

import oracle as orc
from tests import cache_ref as cr

LO, HI = np.array([-1.0, 0.0, -2.0]), np.array([3.0, 1.5, 2.0])


def test_bf16_rounding_known_answers():
    one = 1.0
    cases = {
        one: one,
        one + 2.0 ** -8: one,                      # tie -> even (1.0)
        one + 3 * 2.0 ** -8: one + 2.0 ** -6,      # tie -> even (mantissa ...10)
        one + 2.0 ** -8 + 2.0 ** -20: one + 2.0 ** -7,
        -(one + 2.0 ** -8 + 2.0 ** -20): -(one + 2.0 ** -7),
        0.0: 0.0,
        65280.0: 65280.0,                          # exactly representable (0x477F)
    }
    for x, want in cases.items():
        got = cr.bf16(np.float32(x))
        assert got[()] == np.float32(want), (x, got, want)


def test_encoding_matches_oracle():
    q = cr.random_queries(64, LO, HI, 1)
    enc = cr.encode(q, LO, HI)
    lo32 = LO.astype(np.float32)
    inv = np.float32(1.0) / (HI.astype(np.float32) - lo32)
    for i in range(q.shape[0]):
        p01 = ((q[i, :3] - lo32) * inv).astype(np.float32)
        ref = orc.cache_encode(p01, q[i, 3:6], q[i, 6:9], q[i, 9:12])
        np.testing.assert_allclose(enc[i], ref, rtol=0, atol=1e-12)


def test_fp64_mode_equals_oracle_forward_and_gradient():
    p = cr.random_params(3)
    q = cr.random_queries(200, LO, HI, 2)
    enc = cr.encode(q, LO, HI)
    z = cr.forward_bf16(p, enc, bf16_on=False)
    y = orc.mlp_forward(p.astype(np.float64), enc, 81, 64, 4, 3)
    np.testing.assert_allclose(z, np.log(y), rtol=0, atol=1e-11)
    t = np.exp(np.random.default_rng(4).normal(size=(200, 3))) * 0.5
    loss, g = cr.loss_grad_bf16(p, enc, t, bf16_on=False)
    loss_o, g_o = orc.mlp_loss_grad(p.astype(np.float64), enc, t, 81, 64, 4, 3, rel=True, eps=1e-2)
    assert abs(loss - loss_o) <= 1e-12 * abs(loss_o)
    np.testing.assert_allclose(g, g_o, rtol=1e-9, atol=1e-15)


def test_bf16_model_error_is_small_and_a_transposed_tile_is_not():
    p = cr.random_params(5)
    q = cr.random_queries(2048, LO, HI, 1)
    enc = cr.encode(q, LO, HI)
    z64 = cr.forward_bf16(p, enc, bf16_on=False)
    zb = cr.forward_bf16(p, enc)
    err = np.abs(zb - z64)
    assert 0.0 < np.median(err) < 0.005 and err.max() < 0.03
    # planted mistake: layer 2's 16x16 tile (0, 1) transposed
    bad = p.copy()
    W = cr.weights(bad, 2).copy()
    W[0:16, 16:32] = W[0:16, 16:32].T
    bad[cr.OFF_W[2]:cr.OFF_W[2] + 64 * 64] = W.reshape(-1)
    assert np.abs(cr.forward_bf16(bad, enc) - z64).max() > 0.1
    # gradients: BF16 within a few percent of FP64 per block
    t = np.exp(np.random.default_rng(9).normal(size=(2048, 3))) * 0.5
    _, gb = cr.loss_grad_bf16(p, enc, t)
    _, g64 = cr.loss_grad_bf16(p, enc, t, bf16_on=False)
    for l in range(5):
        for s in (slice(cr.OFF_W[l], cr.OFF_B[l]), slice(cr.OFF_B[l], cr.OFF_B[l] + cr.FAN_OUT[l])):
            rel = np.linalg.norm(gb[s] - g64[s]) / np.linalg.norm(g64[s])
            assert rel < 0.06, (l, s, rel)
