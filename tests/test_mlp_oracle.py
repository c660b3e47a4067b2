"""Pins of the NumPy learned-sensor restatement (oracle/mlp_sensor.py):
finite-difference gradients, the GELU derivative, the plain Jameson flags
against the C oracle's sensor, the feature definitions against direct
evaluation, and the flag rule."""
import numpy as np
import pytest

import mlp_sensor as M
import oracle as O
from pyocfd import cases


def test_gelu_grad_matches_finite_difference():
    x = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(x)).max() < 1e-8
    # closed forms: gelu(0) = 0, gelu'(0) = 1/2, gelu(x) -> x for large x
    assert M.gelu(np.array(0.0)) == 0.0 and abs(M.gelu_grad(np.array(0.0)) - 0.5) < 1e-15
    assert abs(M.gelu(np.array(8.0)) - 8.0) < 1e-12


def test_backward_matches_finite_difference():
    rng = np.random.default_rng(3)
    p = M.init_params(seed=5, dtype=np.float64)
    X = rng.standard_normal((64, M.NF))
    y = (rng.random(64) > 0.6).astype(np.float64)
    loss, g = M.loss_and_grads(p, X, y, wpos=2.0)
    h = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = M.loss_and_grads(p, X, y, wpos=2.0)
            flat[idx] = old - h
            lm, _ = M.loss_and_grads(p, X, y, wpos=2.0)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k].reshape(-1)[idx]) < 1e-6 * max(1.0, abs(fd)), (k, idx)


def test_adam_first_step_is_lr_times_sign():
    # with zero moments, the first Adam step moves each parameter by lr*sign(g)
    p = {"w": np.array([1.0, -2.0, 3.0])}
    g = {"w": np.array([0.5, -4.0, 1e-3])}
    m = {"w": np.zeros(3)}
    v = {"w": np.zeros(3)}
    M.adam_step(p, g, m, v, 1, lr=0.1, eps=0.0)
    assert np.allclose(p["w"], [0.9, -1.9, 2.9], rtol=0, atol=1e-15)


def test_jameson_flags_match_c_oracle():
    case = cases.ramp(40, 24, 10, nsteps=2)
    o = case.load(O.Oracle(case.cfg))
    o.rhs(1)
    fx, fy, fz = o.get_flags()
    # the sensor's pressure with one ghost layer from the oracle's padded field
    kp = M.jameson_kp(o.get_pressure_lap(1), *case.cfg.hy_dp_intv)
    # interior node flags: x flags at face index i+1 hold node i
    assert np.array_equal(kp, fx[:, :, 1:])
    assert np.array_equal(kp, fy[:, 1:, :])
    assert np.array_equal(kp, fz[1:, :, :])
    assert (kp == 1).any() and (kp > 1).any()  # both decisions occur on this flow


def test_feature_definitions_by_direct_evaluation():
    rng = np.random.default_rng(11)
    n = 12
    p = 1 + 0.1 * rng.random(n + 8)
    rho = 1 + 0.1 * rng.random(n + 8)
    U = rng.standard_normal(n + 8)
    c = 0.5 + rng.random(n + 8)
    node = 5
    f = M.features_line(p, rho, U, c, node)
    q = node + 4
    assert f[0] == (p[q] - p[q - 1]) / (p[q] + p[q - 1])
    assert f[1] == (p[q + 1] - p[q]) / (p[q + 1] + p[q])
    assert f[4] == (rho[q + 1] - rho[q]) / (rho[q + 1] + rho[q])
    outer = [abs((p[q + m + 1] - p[q + m]) / (p[q + m + 1] + p[q + m])) for m in (-3, -2, 2, 3)]
    assert f[6] == max(outer)
    assert f[7] == (U[q + 1] - U[q]) / c[q]
    # a uniform state has all-zero features
    one = np.ones(n + 8)
    assert np.all(M.features_line(one, one, one, one, node) == 0)


def test_flag_rule():
    kp = np.array([1, 2, 3, 1, 2, 3], dtype=np.int8)
    prob = np.array([0.1, 0.1, 0.5, 0.9, 0.9, 0.95])
    out = M.flag_rule(prob, kp)
    assert list(out) == [1, 1, 3, 2, 2, 3]
