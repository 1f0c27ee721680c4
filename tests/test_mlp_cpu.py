"""CPU checks of the numpy restatement of the learned Kessler stand-in
(oracle/mlp_ref.py), the reference the GPU's forward pass and gradients are
compared with: its backward pass against central finite differences of the
float64 network (a dropped term, a wrong transpose or a wrong GELU derivative
fails), its parameter count, and BF16 rounding to nearest-even."""
import math

import numpy as np

import sys, os
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref as M  # noqa: E402


def test_parameter_count():
    assert M.NPARAM == 18180 == 8 * 128 + 128 + 128 * 128 + 128 + 128 * 4 + 4


def test_bf16_rounding():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0           # tie -> even
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert M.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    x = np.linspace(-3, 3, 1001)
    assert np.max(np.abs(M.bf16(x) - x) / np.maximum(np.abs(x), 1e-30)) <= 2.0 ** -8


def test_gelu_grad_matches_finite_difference():
    z = np.linspace(-4, 4, 81)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(M.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = M.init_params(1)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.1, M.NH)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.1, M.NH)
    p[M.OFF_B3:] = rng.normal(0, 0.1, M.NO)
    x = rng.normal(0, 1, (16, M.NI))
    t = rng.normal(0, 1, (16, M.NO))
    loss, g = M.loss_and_grad(p, x, t, bf=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) for a, b in
                          ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
                           (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3))] + [np.arange(M.OFF_B3, M.NPARAM)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, t, bf=False)[0] - M.loss_and_grad(pm, x, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
