"""Prefilter MLP on the GPU (bf16 MFMA kernels) against the numpy reference
(tests/mlp_ref.py, pinned by finite differences): forward logits, the loss
gradient, Adam training and the device-resident paths."""
import numpy as np
import pytest

import saber_amd as S
from saber_amd import prefilter as P
from tests import mlp_ref

pytestmark = pytest.mark.gpu


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P.NPARAMS, np.float32)
    p[P.W1:P.B1] = rng.uniform(-1, 1, 64 * 256) * np.sqrt(6 / 64) * scale
    p[P.B1:P.W2] = rng.normal(0, 0.1, 256)
    p[P.W2:P.B2] = rng.uniform(-1, 1, 256 * 256) * np.sqrt(6 / 256) * scale
    p[P.B2:P.W3] = rng.normal(0, 0.1, 256)
    p[P.W3:P.B3] = rng.uniform(-1, 1, 256) * np.sqrt(6 / 256)
    p[P.B3] = 0.3
    return p


@pytest.fixture(scope="module")
def ctx():
    c = S.Context()
    yield c
    c.close()


def test_forward_matches_reference(ctx):
    p = rand_params(1)
    m = P.Mlp(ctx, p)
    rng = np.random.default_rng(2)
    # 64-row groups per wave, 8 waves per block, one block per CU: the sizes
    # cover partial tiles and groups, and (300007) waves that loop over
    # several groups with the next group's rows loaded ahead
    for n in (1, 15, 16, 17, 63, 64, 65, 1000, 4099, 300007):
        x = rng.normal(0, 1, (n, 64)).astype(np.float32)
        got = m.forward(x)
        want, _ = mlp_ref.forward(p, x)
        # fp32 accumulation order differs; a bf16 rounding of a layer-1
        # activation can then land one ulp apart
        err = np.abs(got - want)
        assert np.all(err <= 2e-2 + 1e-2 * np.abs(want)), (n, err.max())
        assert np.mean(err) < 2e-3 * (1 + np.mean(np.abs(want)))
    m.close()


def test_forward_deterministic_at_scale(ctx):
    """4 M rows (every CU loops over many groups) through the forward kernel
    three times: bit-identical logits each time, and a sample within the
    reference tolerance."""
    p = rand_params(4)
    m = P.Mlp(ctx, p)
    x = np.random.default_rng(5).normal(0, 1, (4_000_000, 64)).astype(np.float32)
    first = m.forward(x)
    for _ in range(2):
        assert np.array_equal(m.forward(x), first)
    idx = np.random.default_rng(6).choice(len(x), 20000, replace=False)
    want, _ = mlp_ref.forward(p, x[idx])
    assert np.all(np.abs(first[idx] - want) <= 2e-2 + 1e-2 * np.abs(want))
    m.close()


def test_init_and_params_roundtrip(ctx):
    m = P.Mlp(ctx, seed=7)
    p = m.params()
    assert p.shape == (P.NPARAMS,)
    w1, b1, w2, b2, w3, b3 = P.unpack(p)
    assert np.all(b1 == 0) and np.all(b2 == 0) and b3 == 0
    assert 0.25 < np.abs(w1).max() <= np.sqrt(6 / 64) and np.abs(w2).max() <= np.sqrt(6 / 256)
    q = rand_params(3)
    m.set_params(q)
    assert np.array_equal(m.params(), q)
    m.close()


def test_gradient_matches_reference(ctx):
    p = rand_params(4)
    m = P.Mlp(ctx, p)
    rng = np.random.default_rng(5)
    n = 3000
    x = rng.normal(0, 1, (n, 64)).astype(np.float32)
    t = (rng.random(n) < 0.2).astype(np.float32)
    g, loss = m.backward(x, t, pos_weight=4.0)
    want_loss, want = mlp_ref.loss_and_grad(p, x, t, pos_weight=4.0)
    assert abs(loss - want_loss) <= 1e-3 * abs(want_loss)
    for name, a, b in (("W1", P.W1, P.B1), ("b1", P.B1, P.W2), ("W2", P.W2, P.B2), ("b2", P.B2, P.W3),
                       ("w3", P.W3, P.B3), ("b3", P.B3, P.NPARAMS)):
        gg, ww = g[a:b].astype(np.float64), want[a:b]
        cos = gg @ ww / (np.linalg.norm(gg) * np.linalg.norm(ww) + 1e-30)
        rel = abs(np.linalg.norm(gg) - np.linalg.norm(ww)) / (np.linalg.norm(ww) + 1e-30)
        assert cos > 0.999 and rel < 0.02, (name, cos, rel)
    # parameters are unchanged by backward
    assert np.array_equal(m.params(), p)
    m.close()


def test_training_learns_a_nonlinear_rule(ctx):
    rng = np.random.default_rng(8)
    n = 16384
    x = rng.normal(0, 1, (n, 64)).astype(np.float32)
    t = (x[:, 0] * x[:, 1] > 0).astype(np.float32)   # sign XOR: not linearly separable
    m = P.Mlp(ctx, seed=9)
    losses = [m.train_step(x[i * 2048:(i + 1) * 2048], t[i * 2048:(i + 1) * 2048], lr=3e-3)
              for _ in range(30) for i in range(8)]
    assert losses[-1] < 0.5 * losses[0]
    xt = rng.normal(0, 1, (4096, 64)).astype(np.float32)
    tt = xt[:, 0] * xt[:, 1] > 0
    acc = np.mean((m.forward(xt) > 0) == tt)
    assert acc > 0.9, acc
    m.close()


def test_resident_paths_equal_host_paths(ctx):
    p = rand_params(11)
    rng = np.random.default_rng(12)
    n = 5000
    x = rng.normal(0, 1, (n, 64)).astype(np.float32)
    t = (rng.random(n) < 0.3).astype(np.float32)
    a = P.Mlp(ctx, p)
    b = P.Mlp(ctx, p)
    b.upload(x, t)
    assert np.array_equal(a.forward(x[1000:3000]), b.forward_resident(1000, 2000))
    la = a.train_step(x[:4000], t[:4000], lr=1e-3, pos_weight=2.0)
    lb = b.train_resident(0, 4000, lr=1e-3, pos_weight=2.0)
    # gradient reductions add with atomics, so the two runs agree to rounding
    assert abs(la - lb) <= 1e-5 * abs(la)
    pa, pb = a.params(), b.params()
    assert np.allclose(pa, pb, rtol=0, atol=1e-5)
    f_ms, s_ms = b.time(4000, warmup=1, iters=2)
    assert f_ms > 0 and s_ms > 0
    assert np.array_equal(b.params(), pb)   # timing restores the state
    with pytest.raises(S.HipError):
        b.forward_resident(4000, 2000)
    a.close()
    b.close()
