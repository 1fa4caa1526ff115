"""GPU: the learned-screen network's kernels against the CPU oracle
(oracle/mlp_oracle.py), its features, and screened builds.

Tolerances: losses 1e-5 relative; gradients max|dg| <= 1e-3 max|g| (FP32
sums in another order can move a hidden activation across a bf16 rounding
boundary); one Adam step within 1e-3 lr of the oracle on entries whose
gradient is not noise; network outputs 1e-4 (per-lane FP32 evaluator) and
5e-3 (bf16 matrix-core evaluator, median 1e-5).
"""
import os

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from tlno_amd import configs, learned
from tlno_amd.fock import DirectPlan, ScreeningOptions

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = np.concatenate([rng.normal(-1.0, 1.5, (n, 20)), np.eye(6)[rng.integers(0, 6, n)],
                        np.eye(6)[rng.integers(0, 6, n)]], axis=1).astype(np.float32)
    t = (0.3 * X[:, 0] - 0.2 * X[:, 5] + 0.5 * np.sin(X[:, 3]) - 1.0).astype(np.float32)
    return X, t


def _net(seed=3, n=512):
    X, t = _data(n, seed)
    net = learned.ScreenNetwork(seed=seed)
    net.set_normalization(X)
    return net, X, t


@pytest.mark.parametrize("w_under", [1.0, 4.0])
def test_gradient_matches_oracle(w_under):
    net, X, t = _net()
    loss, g = net.loss_grad(X, t, w_under)
    lo, go = mo.loss_grad(net.params.astype(np.float64), X, t, w_under)
    print(f"loss {loss:.6e} oracle {lo:.6e}; max|dg| {np.abs(g - go).max():.2e} of max|g| {np.abs(go).max():.2e}")
    assert loss == pytest.approx(lo, rel=1e-5)
    assert np.abs(g - go).max() <= 1e-3 * np.abs(go).max() + 1e-7
    assert np.dot(g, go) / (np.linalg.norm(g) * np.linalg.norm(go)) > 0.999999


def test_adam_step_matches_oracle():
    net, X, t = _net(seed=5)
    p0 = net.params.astype(np.float64)
    _, go = mo.loss_grad(p0, X, t, 4.0)
    lr = 1e-3
    net.adam_step(X, t, lr, 4.0)
    p1 = net.params.astype(np.float64)
    want, _, _ = mo.adam_step(p0, go, np.zeros(mo.NPARAM), np.zeros(mo.NPARAM), 1, lr)
    sig = np.abs(go) > 1e-3 * np.abs(go).max()
    d = np.abs(p1[:mo.NPARAM] - want[:mo.NPARAM])
    assert sig.sum() > 1000
    assert d[sig].max() <= 1e-3 * lr + 1e-7
    assert d.max() <= 2 * lr + 1e-7
    assert np.array_equal(p1[mo.NPARAM:], p0[mo.NPARAM:])


def test_predict_lane_and_matrix_cores_match_oracle():
    net, X, _ = _net(seed=7, n=1000)   # not a multiple of 64: the last wave is partial
    yo, _ = mo.forward(net.params.astype(np.float64), X)
    y_lane = net.predict(X, matrix_cores=False)
    y_mfma = net.predict(X, matrix_cores=True)
    dl, dm = np.abs(y_lane - yo), np.abs(y_mfma - yo)
    print(f"lane max {dl.max():.2e}; matrix cores max {dm.max():.2e} median {np.median(dm):.2e}")
    assert dl.max() <= 1e-4 * (1 + np.abs(yo).max())
    assert dm.max() <= 5e-3 * (1 + np.abs(yo).max())
    assert np.median(dm) <= 1e-5 * (1 + np.abs(yo).max())


def test_training_reduces_loss():
    X, t = _data(20000, 11)
    net = learned.ScreenNetwork(seed=1)
    net.set_normalization(X)
    l0, _ = net.loss_grad(X, t)
    hist = net.train(X, t, epochs=4, batch=512, lr=3e-3, seed=2)
    l1, _ = net.loss_grad(X, t)
    print(f"loss {l0:.4f} -> {l1:.4f}; epochs {hist}")
    assert l1 < 0.1 * l0
    assert hist[-1] < hist[0]


def test_features_consistent_with_plan():
    b, _ = configs.water_cluster(3)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    ga, gb, Qb, cb = plan.pairs()
    cls = learned.pair_classes(plan)
    rng = np.random.default_rng(0)
    P, Q = learned.sample_quartets(plan, 3000, rng)
    F = learned.quartet_features(plan, P, Q)
    FT = learned.quartet_features(plan, Q, P)
    # pair blocks swap with the quartet's sides; quartet distances are symmetric
    np.testing.assert_array_equal(F[:, 0:8], FT[:, 8:16])
    np.testing.assert_allclose(F[:, 17:20], FT[:, 17:20], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(F[:, 0], np.log10(Qb[P]), rtol=0, atol=2e-6 * np.abs(np.log10(Qb[P])).max())
    np.testing.assert_allclose(F[:, 16], np.log10(Qb[P] * Qb[Q]), atol=1e-5)
    assert np.array_equal(np.argmax(F[:, 20:26], 1), cls[P]) and np.array_equal(np.argmax(F[:, 26:32], 1), cls[Q])
    assert np.all(F[:, 20:32].sum(1) == 2.0)
    # |AB| and p_min from the basis itself
    centers = {}
    exps = {}
    for g, comps in enumerate(plan.group_components()):
        i = int(comps[0])
        centers[g] = b.centers[i]
        o = int(np.sum(b.nprim[:i]))
        exps[g] = b.exps[o:o + b.nprim[i]]
    ab = np.array([np.linalg.norm(centers[ga[p]] - centers[gb[p]]) for p in P])
    np.testing.assert_allclose(F[:, 4], ab, rtol=1e-6, atol=1e-6)
    pmin = np.array([exps[ga[p]].min() + exps[gb[p]].min() for p in P])
    np.testing.assert_allclose(F[:, 1], np.log10(pmin), rtol=1e-6, atol=1e-6)
    # the label is <= 0 (Cauchy-Schwarz) and reaches 0 on diagonal quartets
    t = learned.labels(plan, P, Q)
    assert t.max() <= 1e-9
    d = np.arange(0, plan.profile.pair_groups, 7)
    assert np.abs(learned.labels(plan, d, d)).max() < 1e-9


def _weights():
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights in the package")
    return learned.load_weights()


@pytest.mark.parametrize("nwat,tau", [(10, 3e-10), (20, 1e-9)])
def test_learned_screen_gain_and_error(nwat, tau):
    params, margin = _weights()
    b, D = configs.water_cluster(nwat)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J0, K0 = plan.jk(D)
    Jd, Kd = plan.jk(D, density_threshold=tau)
    sq_d, pq_d = plan.last_build_stats()
    rd = plan.last_route_stats()
    plan.set_learned_screen(params, margin)
    Jl, Kl = plan.jk(D, density_threshold=tau)
    sq_l, pq_l = plan.last_build_stats()
    rl = plan.last_route_stats()
    plan.set_learned_screen(None)
    e_d = max(np.abs(Jd - J0).max(), np.abs(Kd - K0).max())
    e_l = max(np.abs(Jl - J0).max(), np.abs(Kl - K0).max())
    print(f"water{nwat} tau {tau:g}: density-only {sq_d} quartets err {e_d:.2e}; learned {sq_l} quartets "
          f"({sq_l / sq_d:.3f}) err {e_l:.2e}; routes {rd} {rl}")
    assert rl["candidates"] == rd["candidates"] and rl["density_kept"] == rd["density_kept"]
    assert sq_l < sq_d and pq_l < pq_d      # the gain over the method without the network
    assert sq_l == rl["kept"]
    assert e_l <= 1e-7
    # off again: the density-only build returns
    J2, K2 = plan.jk(D, density_threshold=tau)
    assert plan.last_build_stats() == (sq_d, pq_d)
    assert np.abs(J2 - Jd).max() < 1e-13 and np.abs(K2 - Kd).max() < 1e-13


def test_learned_screen_without_density_uses_schwarz_threshold():
    params, margin = _weights()
    b, D = configs.water_cluster(6)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-10), cartesian=True)
    J0, K0 = plan.jk(D)
    sq0, _ = plan.last_build_stats()
    plan.set_learned_screen(params, margin)
    J1, K1 = plan.jk(D)
    sq1, _ = plan.last_build_stats()
    r = plan.last_route_stats()
    assert r["candidates"] == sq0 and r["density_kept"] == sq0
    assert sq1 <= sq0
    assert max(np.abs(J1 - J0).max(), np.abs(K1 - K0).max()) < 1e-7
