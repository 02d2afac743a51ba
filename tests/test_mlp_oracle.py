"""The CPU restatement of the surrogate network (oracle/mlp_ref.py), pinned
by finite differences, the closed forms of GELU and the Adam update."""
import numpy as np
import pytest
from scipy.special import erf

from oracle import mlp_ref as M


def test_bf16_rounding():
    # exactly representable values stay; ties go to even; 1 + 2^-8 rounds to 1
    x = np.array([1.0, 1.5, -2.25, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 3.0e38, 1.0e-30], dtype=np.float32)
    r = M.bf16(x)
    assert list(r[:3]) == [1.0, 1.5, -2.25]
    assert r[3] == 1.0 and r[4] == 1 + 2 ** -6
    assert abs(r[5] - 3.0e38) / 3.0e38 < 2 ** -8 and abs(r[6] - 1e-30) / 1e-30 < 2 ** -8


def test_gelu_closed_form():
    z = np.linspace(-5, 5, 101)
    np.testing.assert_allclose(M.gelu(z), z * 0.5 * (1 + erf(z / np.sqrt(2))), rtol=1e-14, atol=1e-300)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)
    assert M.gelu(0.0) == 0.0 and M.gelu_grad(0.0) == 0.5


def test_gradient_matches_finite_differences():
    widths = [5, 8, 7, 6]
    rng = np.random.default_rng(3)
    params = rng.standard_normal(sum(o * i + o for i, o in zip(widths[:-1], widths[1:]))) * 0.5
    x = rng.standard_normal((4, 5))
    t = rng.standard_normal((4, 6))
    _, g = M.loss_and_gradient(params, widths, x, t, bf=False)
    h = 1e-6
    fd = np.zeros_like(params)
    for k in range(len(params)):
        e = np.zeros_like(params)
        e[k] = h
        fd[k] = (M.loss_and_gradient(params + e, widths, x, t, bf=False)[0] -
                 M.loss_and_gradient(params - e, widths, x, t, bf=False)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_adam_first_step_is_signed_lr():
    # with m = v = 0, the first bias-corrected step is lr * g / (|g| + eps)
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p, m, v = M.adam_step(np.zeros(4), np.zeros(4), np.zeros(4), g, t=1, lr=1e-3)
    np.testing.assert_allclose(p, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_init_is_seeded_and_bounded():
    widths = [66, 64, 128]
    a = M.init_params(widths, 7)
    b = M.init_params(widths, 7)
    np.testing.assert_array_equal(a, b)
    (W0, b0), (W1, b1) = M.unpack(a, widths)
    assert np.abs(W0).max() <= 1 / np.sqrt(66) and np.abs(W1).max() <= 1 / np.sqrt(64)
    assert abs(W0.mean()) < 0.01 and W0.std() == pytest.approx(1 / np.sqrt(66) / np.sqrt(3), rel=0.05)
