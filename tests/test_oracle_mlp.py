"""Pins of the MLP oracle: GELU known values, gradients against central finite
differences, Adam against its closed form on one step."""
import math

import numpy as np

from oracle import mlp as M


def test_gelu_known_values():
    z = np.array([-3.0, -1.0, 0.0, 1.0, 2.0])
    ref = z * 0.5 * (1 + np.array([math.erf(v / math.sqrt(2)) for v in z]))
    np.testing.assert_allclose(M.gelu(z), ref, rtol=1e-15)
    assert abs(M.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), rtol=1e-6, atol=1e-9)


def test_gradients_match_finite_differences():
    dims = (12, 9, 7, 6)
    params = [p.astype(np.float64) for p in M.init_params(3, dims)]
    rng = np.random.default_rng(4)
    params[1] = rng.standard_normal(params[1].shape) * 0.1
    params[3] = rng.standard_normal(params[3].shape) * 0.1
    x = rng.standard_normal((5, 12))
    y = rng.integers(0, 5, 5)
    r = rng.standard_normal(5)
    _, grads = M.loss_and_grads(params, x, y, r)
    h = 1e-6
    for pi in range(6):
        flat = params[pi].reshape(-1)
        for idx in rng.choice(flat.size, min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = M.loss_and_grads(params, x, y, r)
            flat[idx] = old - h
            lm, _ = M.loss_and_grads(params, x, y, r)
            flat[idx] = old
            num = (lp - lm) / (2 * h)
            assert abs(grads[pi].reshape(-1)[idx] - num) <= 1e-6 + 1e-5 * abs(num), (pi, idx)


def test_adam_first_step_closed_form():
    # first step: m = 0.1 g, v = 0.001 g^2 -> update = lr * g / (|g| + eps) ~ lr * sign(g)
    p = [np.array([1.0, -2.0, 0.5])]
    g = [np.array([0.3, -0.2, 0.0])]
    out, m, v = M.adam_step(p, g, [np.zeros(3)], [np.zeros(3)], 1, lr=1e-3)
    np.testing.assert_allclose(out[0], [1.0 - 1e-3 * 0.3 / (0.3 + 1e-8), -2.0 + 1e-3 * 0.2 / (0.2 + 1e-8), 0.5])


def test_decide_confidence_fallback():
    o = np.array([[5.0, 0, 0, 0, 0, 1.0], [0.1, 0.2, 0.0, 0.1, 0.0, 0.0]])
    np.testing.assert_array_equal(M.decide(o), [2, 16])
