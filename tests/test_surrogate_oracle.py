"""Pins the surrogate MLP oracle (oracle/surrogate.py): analytic gradients vs
central finite differences (float64), Adam's first step vs its closed form,
Spearman vs scipy.stats.spearmanr, and the BF16 rounding helper on known
values."""
import numpy as np
import pytest
import scipy.stats

import surrogate as SO


def _flat(p):
    return np.concatenate([p[k].ravel() for k in sorted(p)])


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7, -2.5, 3.0e38, 0.0])
    r = SO.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                        # tie -> even
    assert r[2] == 1 + 2 ** -7                # above the tie -> up
    assert r[3] == 1 + 2 ** -7                # representable
    assert r[4] == -2.5 and r[6] == 0.0
    assert np.isfinite(r[5])


@pytest.mark.parametrize("round_bf16", [False])
def test_gradients_against_finite_differences(round_bf16):
    rng = np.random.default_rng(0)
    D, B = 6, 9
    p = SO.init_params(D, seed=1)
    p["b1"] = rng.normal(0, 0.1, SO.HIDDEN)
    p["b2"] = rng.normal(0, 0.1, SO.HIDDEN)
    p["b3"] = np.array([0.3])
    X = rng.uniform(-100, 100, (B, D))
    t = rng.normal(0, 1, B)
    L0, g = SO.loss_and_grads(p, X, t)
    h = 1e-6
    for name, idx in [("W1", (3, 2)), ("W1", (200, 5)), ("b1", (17,)), ("W2", (5, 250)),
                      ("W2", (128, 3)), ("b2", (99,)), ("w3", (7,)), ("w3", (255,)),
                      ("b3", (0,))]:
        q = {k: v.copy() for k, v in p.items()}
        q[name][idx] += h
        Lp, _ = SO.loss_and_grads(q, X, t)
        q[name][idx] -= 2 * h
        Lm, _ = SO.loss_and_grads(q, X, t)
        fd = (Lp - Lm) / (2 * h)
        assert g[name][idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), name


def test_adam_first_step_closed_form():
    p = {"a": np.array([1.0, -2.0, 3.0])}
    g = {"a": np.array([0.5, -4.0, 1e-3])}
    m = {"a": np.zeros(3)}
    v = {"a": np.zeros(3)}
    p2, m2, v2 = SO.adam_step(p, g, m, v, 1, lr=0.01, eps=0.0)
    # m_hat = g, v_hat = g^2 -> step = -lr * sign(g)
    np.testing.assert_allclose(p2["a"], p["a"] - 0.01 * np.sign(g["a"]), rtol=1e-12)
    np.testing.assert_allclose(m2["a"], 0.1 * g["a"])
    np.testing.assert_allclose(v2["a"], 0.001 * g["a"] ** 2)


def test_spearman_matches_scipy():
    rng = np.random.default_rng(3)
    for n in (5, 50, 1000):
        a = rng.normal(size=n)
        b = a + rng.normal(scale=0.7, size=n)
        a[::7] = a[0]                          # ties
        b[::5] = 0.25
        assert SO.spearman(a, b) == pytest.approx(scipy.stats.spearmanr(a, b)[0], abs=1e-12)


def test_training_fits_a_smooth_function():
    # a few hundred Adam steps fit log(1 + |x|^2) well enough to rank it
    rng = np.random.default_rng(5)
    D = 4
    p = SO.init_params(D, seed=2)
    m = {k: np.zeros_like(v) for k, v in p.items()}
    v = {k: np.zeros_like(x) for k, x in p.items()}
    X = rng.uniform(-100, 100, (512, D))
    y = np.log1p((X ** 2).sum(1))
    mu, sd = y.mean(), y.std()
    for step in range(1, 301):
        idx = rng.integers(0, 512, 64)
        _, g = SO.loss_and_grads(p, X[idx], (y[idx] - mu) / sd)
        p, m, v = SO.adam_step(p, g, m, v, step, lr=3e-3)
    Xt = rng.uniform(-100, 100, (256, D))
    pred, _ = SO.forward(p, Xt)
    assert SO.spearman(pred, np.log1p((Xt ** 2).sum(1))) > 0.9
