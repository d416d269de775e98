"""The numpy MLP reference (used by the GPU prefilter tests) pinned by finite
differences: a dropped term, a transposed weight or a wrong ReLU mask in the
gradient formulas fails here."""
import numpy as np

from saber_amd.prefilter import NPARAMS, bf16_bits, bf16_float
from tests import mlp_ref


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.0, 3.14159265, 65504.0, 0.0], np.float32)
    got = bf16_float(bf16_bits(x))
    # ties go to even: 1 + 2^-8 -> 1.0, 1 + 3*2^-8 -> 1 + 2^-6
    assert got.tolist() == [1.0, 1.0, 1.015625, -2.0, 3.140625, 65536.0, 0.0]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.2, NPARAMS)
    x = rng.normal(0, 1, (6, 64))
    t = (rng.random(6) > 0.5).astype(np.float64)
    loss, g = mlp_ref.loss_and_grad(p, x, t, pos_weight=3.0, bf16=False)
    idx = np.concatenate([rng.choice(NPARAMS - 1, 40, replace=False), [NPARAMS - 1]])
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = mlp_ref.loss_and_grad(q, x, t, pos_weight=3.0, bf16=False)
        q[i] -= 2 * eps
        lm, _ = mlp_ref.loss_and_grad(q, x, t, pos_weight=3.0, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_at_zero_logits():
    # all-zero parameters give y = 0: loss = log 2 per row (weighted)
    p = np.zeros(NPARAMS)
    x = np.ones((4, 64))
    t = np.array([1.0, 0.0, 1.0, 0.0])
    loss, g = mlp_ref.loss_and_grad(p, x, t, pos_weight=2.0, bf16=False)
    assert abs(loss - np.log(2) * (2 + 1 + 2 + 1) / 4) < 1e-12
