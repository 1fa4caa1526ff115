"""CPU: pins the learned-screen oracle (oracle/mlp_oracle.py).

The network has no reference counterpart, so the oracle is pinned by what
the mathematics fixes: central finite differences of the loss (rounding off),
Adam's closed-form first step (bias correction makes it -lr g / (|g| + eps)),
the two-step recursion written out by hand, exact bfloat16 rounding cases,
and a network with zero weights whose output is its output bias.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _params(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(mo.NALL)
    w[mo.OFF_W1:mo.OFF_B1] = rng.uniform(-1, 1, mo.NH * mo.NIN) * np.sqrt(6 / 32)
    w[mo.OFF_B1:mo.OFF_W2] = rng.uniform(-0.1, 0.1, mo.NH)
    w[mo.OFF_W2:mo.OFF_B2] = rng.uniform(-1, 1, mo.NH * mo.NH) * np.sqrt(6 / 64)
    w[mo.OFF_B2:mo.OFF_W3] = rng.uniform(-0.1, 0.1, mo.NH)
    w[mo.OFF_W3:mo.OFF_B3] = rng.uniform(-1, 1, mo.NH) * 0.3
    w[mo.OFF_B3] = -0.5
    w[mo.OFF_MEAN:mo.OFF_ISTD] = rng.normal(0, 1, mo.NIN)
    w[mo.OFF_ISTD:] = rng.uniform(0.5, 2.0, mo.NIN)
    return w


def test_bf16_exact_cases():
    e = 2.0 ** -8
    vals = np.array([1.0, 1 + e, 1 + 3 * e, 1 + e + 2.0 ** -20, -(1 + 3 * e), 0.0, 3.0e38, 2.0 ** -130])
    want = np.array([1.0, 1.0, 1 + 4 * e, 1 + 2 * e, -(1 + 4 * e), 0.0, np.float32(3.0e38), 2.0 ** -130])
    got = mo.bf16(vals)
    assert got[:6].tolist() == want[:6].tolist()
    assert abs(got[6] / 3.0e38 - 1) < 2 ** -8       # relative error of a bf16 rounding
    assert got[7] == 2.0 ** -130                     # a power of two is exact


@pytest.mark.parametrize("w_under", [1.0, 4.0])
def test_gradient_matches_finite_differences(w_under):
    w = _params(1)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1.5, (48, mo.NIN))
    t = rng.normal(-1.0, 1.0, 48)
    _, g = mo.loss_grad(w, X, t, w_under, rounding=False)
    idx = np.concatenate([rng.choice(np.arange(mo.OFF_B1), 12, replace=False),
                          mo.OFF_B1 + rng.choice(mo.NH, 4, replace=False),
                          mo.OFF_W2 + rng.choice(mo.NH * mo.NH, 12, replace=False),
                          mo.OFF_B2 + rng.choice(mo.NH, 4, replace=False),
                          mo.OFF_W3 + rng.choice(mo.NH, 6, replace=False), [mo.OFF_B3]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mo.loss_grad(wp, X, t, w_under, False)[0] - mo.loss_grad(wm, X, t, w_under, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_zero_network_is_its_bias():
    w = np.zeros(mo.NALL)
    w[mo.OFF_ISTD:] = 1.0
    w[mo.OFF_B3] = -0.75
    X = np.random.default_rng(3).normal(size=(10, mo.NIN))
    t = np.linspace(-2, 1, 10)
    y, _ = mo.forward(w, X)
    assert np.all(y == -0.75)
    loss, g = mo.loss_grad(w, X, t, 4.0)
    r = t + 0.75
    assert loss == pytest.approx(np.mean(np.where(r > 0, 4.0, 1.0) * r * r), rel=1e-14)
    assert g[mo.OFF_B3] == pytest.approx(np.mean(-2 * np.where(r > 0, 4.0, 1.0) * r), rel=1e-14)


def test_adam_first_step_closed_form():
    w0 = _params(4)
    g = np.random.default_rng(5).normal(0, 1e-3, mo.NPARAM)
    w1, m, v = mo.adam_step(w0, g, np.zeros(mo.NPARAM), np.zeros(mo.NPARAM), 1, 1e-3)
    np.testing.assert_allclose(w1[:mo.NPARAM] - w0[:mo.NPARAM], -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-18)
    assert np.array_equal(w1[mo.NPARAM:], w0[mo.NPARAM:])  # normalization is not trained


def test_adam_two_steps_by_hand():
    w0 = np.zeros(mo.NALL)
    g1 = np.full(mo.NPARAM, 2.0)
    g2 = np.full(mo.NPARAM, -1.0)
    w1, m, v = mo.adam_step(w0, g1, np.zeros(mo.NPARAM), np.zeros(mo.NPARAM), 1, 0.1)
    w2, _, _ = mo.adam_step(w1, g2, m, v, 2, 0.1)
    # m2 = 0.9*0.1*2 + 0.1*(-1) = 0.08, v2 = 0.999*0.001*4 + 0.001*1 = 0.004996
    mh, vh = 0.08 / (1 - 0.81), 0.004996 / (1 - 0.998001)
    step2 = -0.1 * mh / (np.sqrt(vh) + 1e-8)
    np.testing.assert_allclose(w2[:mo.NPARAM], -0.1 * 2 / (2 + 1e-8) + step2, rtol=1e-12)
