"""GPU parity of the surrogate network kernels (BF16 MFMA products, FP32
accumulation, Adam) against the CPU restatement oracle/mlp_ref.py, which
rounds to BF16 at the same points and sums in float64.

Tolerances: the GPU sums in float32 in another order, so a value rounded to
BF16 can land one BF16 step (2^-8 relative) away from the oracle's; that
moves outputs and gradients by about 1e-3 relative.  Forward: rel-L2 5e-3;
gradients: rel-L2 2e-2 per layer; Adam (pure FP32): 1e-6.
"""
import numpy as np
import pytest

import se2dr_fork_amd as S
from conftest import rel_l2
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu

SMALL = [66, 128, 128, 192]
FULL = [66, 1024, 1024, 1024, 4096]  # the surrogate's network


def layer_slices(widths):
    off = 0
    for i, o in zip(widths[:-1], widths[1:]):
        yield slice(off, off + o * i), slice(off + o * i, off + o * i + o)
        off += o * i + o


def test_init_matches_oracle():
    m = S.Mlp(SMALL, batch=64, seed=11)
    np.testing.assert_array_equal(m.params(), M.init_params(SMALL, 11).astype(np.float32))
    m.close()


@pytest.mark.parametrize("widths,batch,rows", [(SMALL, 64, 100), (FULL, 256, 300)])
def test_forward_matches_oracle(widths, batch, rows):
    m = S.Mlp(widths, batch=batch, seed=5)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((rows, widths[0])).astype(np.float32)
    y = m.forward(x)
    y0, _ = M.forward(m.params(), widths, x)
    assert rel_l2(y, y0) < 5e-3
    m.close()


@pytest.mark.parametrize("widths,batch", [(SMALL, 64), (FULL, 256)])
def test_gradient_matches_oracle(widths, batch):
    m = S.Mlp(widths, batch=batch, seed=9)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((batch, widths[0])).astype(np.float32)
    t = rng.standard_normal((batch, widths[-1])).astype(np.float32)
    loss, g = m.gradient(x, t)
    loss0, g0 = M.loss_and_gradient(m.params(), widths, x, t)
    assert loss == pytest.approx(loss0, rel=1e-3)
    for sw, sb in layer_slices(widths):
        assert rel_l2(g[sw], g0[sw]) < 2e-2
        assert rel_l2(g[sb], g0[sb]) < 2e-2
    # BF16 storage changes the gradient only a little against exact FP64 arithmetic
    _, g64 = M.loss_and_gradient(m.params(), widths, x, t, bf=False)
    assert rel_l2(g, g64) < 5e-2
    m.close()


def test_adam_step_matches_oracle():
    m = S.Mlp(SMALL, batch=64, seed=3)
    p0 = m.params().astype(np.float64)
    rng = np.random.default_rng(4)
    g1, g2 = rng.standard_normal(m.nparams) * 1e-2, rng.standard_normal(m.nparams) * 1e-2
    m.adam_step(g1, 1e-3)
    m.adam_step(g2, 1e-3)
    p, mm, vv = M.adam_step(p0, 0 * p0, 0 * p0, g1.astype(np.float32).astype(np.float64), 1, 1e-3)
    p, mm, vv = M.adam_step(p, mm, vv, g2.astype(np.float32).astype(np.float64), 2, 1e-3)
    np.testing.assert_allclose(m.params(), p, rtol=1e-5, atol=1e-7)
    m.close()


def test_training_reduces_the_loss():
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (512, 66)).astype(np.float32)
    A = rng.standard_normal((66, 192)) / 8
    t = np.sin(x @ A).astype(np.float32)
    m = S.Mlp(SMALL, batch=64, seed=1)
    m.set_data(x, t)
    losses = m.train(40, lr=1e-3, seed=2)
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.2 * losses[0]
    # the trained network predicts the training rows better than their mean
    y = m.forward(x)
    assert rel_l2(y, t) < 0.5 * rel_l2(np.broadcast_to(t.mean(axis=0), t.shape), t)
    train_ms, fwd_ms = m.timing()
    assert train_ms > 0 and fwd_ms > 0
    m.close()


def test_mlp_argument_errors():
    with pytest.raises(S.Se2wError, match="multiple of 64"):
        S.Mlp([66, 100, 64], batch=64)
    with pytest.raises(S.Se2wError):
        S.Mlp([66, 64, 64], batch=48)
    m = S.Mlp(SMALL, batch=64)
    with pytest.raises(S.Se2wError, match="64 rows"):
        m.gradient(np.zeros((32, 66)), np.zeros((32, 192)))
    with pytest.raises(S.Se2wError, match="training data"):
        m.train(1)
    m.close()
