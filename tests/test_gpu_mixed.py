"""GPU: mixed-precision J/K (FP32 recurrences for quartets with
Q_ab Q_cd max|D| < 1e-6, FP64 roots and FP64 accumulation) against the FP64
GPU build, which is itself checked against the oracle and the reference, and
on water10 against the oracle's FP64 J/K directly.

Tolerance: max-abs J/K error <= 1e-7 (the target for the mixed route); the
RMS error is reported.
"""
import numpy as np
import pytest

from tlno_amd import configs
from tlno_amd.fock import DirectPlan, ScreeningOptions

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nwat", [10, 40])
def test_mixed_precision_error_bound(nwat):
    b, D = configs.water_cluster(nwat)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J0, K0 = plan.jk(D)
    plan.set_mixed_precision(1e-6)
    J1, K1 = plan.jk(D)
    st = plan.last_mixed_stats()
    sq, pq = plan.last_build_stats()
    plan.set_mixed_precision(0.0)
    ej, ek = np.abs(J1 - J0).max(), np.abs(K1 - K0).max()
    rj = np.sqrt(np.mean((J1 - J0) ** 2))
    rk = np.sqrt(np.mean((K1 - K0) ** 2))
    print(f"water{nwat}: FP32 share {st['fp32_shell_quartets'] / sq:.3f} of shell quartets, "
          f"{st['fp32_prim_quartets'] / pq:.3f} of primitive quartets; max|dJ| {ej:.2e} max|dK| {ek:.2e} "
          f"rms {rj:.2e} {rk:.2e}")
    assert st["fp32_shell_quartets"] > 0
    assert ej <= 1e-7 and ek <= 1e-7
    # the gate is off again: bit-identical to the first FP64 build
    J2, K2 = plan.jk(D)
    assert np.array_equal(J2, J0) or np.abs(J2 - J0).max() < 1e-13


def test_mixed_and_learned_routes_against_the_oracle_water10():
    """The mixed route, and learned screening over it, against the CPU
    oracle's FP64 J/K of water10 (tests/golden/water10.npz, Cartesian,
    Schwarz 1e-12): max-abs <= 1e-7 for both routes, <= 1e-10 for the plain
    FP64 GPU build."""
    import os
    from tlno_amd import learned
    from test_oracle import basis_hash
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "water10.npz"))
    b, D = configs.water_cluster(10)
    assert basis_hash(b) == str(g["basis_sha"])
    assert abs(D.sum() - float(g["D_sum"])) <= 1e-12 and abs(np.linalg.norm(D) - float(g["D_fro"])) <= 1e-12
    Jo, Ko = g["J"], g["K"]
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J0, K0 = plan.jk(D)
    e0 = max(np.abs(J0 - Jo).max(), np.abs(K0 - Ko).max())
    plan.set_mixed_precision(1e-6)
    J1, K1 = plan.jk(D)
    e1 = max(np.abs(J1 - Jo).max(), np.abs(K1 - Ko).max())
    r1 = np.sqrt(np.mean(np.concatenate([(J1 - Jo).ravel(), (K1 - Ko).ravel()]) ** 2))
    params, margin = learned.load_weights()
    plan.set_learned_screen(params, margin)
    J2, K2 = plan.jk(D, density_threshold=2e-9)
    e2 = max(np.abs(J2 - Jo).max(), np.abs(K2 - Ko).max())
    plan.set_learned_screen(None)
    plan.set_mixed_precision(0.0)
    print(f"water10 vs oracle: FP64 {e0:.2e}; mixed {e1:.2e} (rms {r1:.2e}); learned+mixed {e2:.2e}")
    assert e0 <= 1e-10
    assert e1 <= 1e-7 and e2 <= 1e-7


def test_mixed_gate_zero_and_huge():
    b, D = configs.water_cluster(4)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J0, K0 = plan.jk(D)
    plan.set_mixed_precision(1e30)  # everything FP32
    J1, K1 = plan.jk(D)
    sq, _ = plan.last_build_stats()
    assert plan.last_mixed_stats()["fp32_shell_quartets"] == sq
    # all-FP32 is still close in relative terms
    assert np.abs(J1 - J0).max() < 1e-4 * np.abs(J0).max()
    with pytest.raises(Exception):
        plan.set_mixed_precision(-1.0)


def test_mixed_precision_mixed_contractions():
    """organic30 (1-6 primitives per shell): the lane evaluator takes
    bucketed list tasks there; the FP64 prefix / FP32 suffix split of those
    lists must cover every quartet once (shell-quartet count equal to the
    FP64 build's) and stay within the mixed-route bound."""
    b, D = configs.organic30()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J0, K0 = plan.jk(D)
    sq0, pq0 = plan.last_build_stats()
    plan.set_mixed_precision(1e-6)
    J1, K1 = plan.jk(D)
    st = plan.last_mixed_stats()
    sq1, pq1 = plan.last_build_stats()
    plan.set_mixed_precision(0.0)
    ej, ek = np.abs(J1 - J0).max(), np.abs(K1 - K0).max()
    print(f"organic30: FP32 share {st['fp32_shell_quartets'] / sq1:.3f}; max|dJ| {ej:.2e} max|dK| {ek:.2e}")
    assert (sq1, pq1) == (sq0, pq0)
    assert 0 < st["fp32_shell_quartets"] < sq1
    assert ej <= 1e-7 and ek <= 1e-7


def test_mixed_precision_after_autotune():
    """The autotuner times every lane class with its FP32 kernels as well; a
    class whose FP32 kernels are not faster runs its FP32 share on the FP64
    kernels.  The mixed build after autotuning must evaluate every quartet
    once (same counts as FP64), keep the mixed-route bound, and the report
    must carry the FP32 timing of every lane class."""
    b, D = configs.water_cluster(10)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    rep = plan.autotune(schwarz=1e-12, repeats=1)
    assert rep and all("lane_fp32_ms" in v and v["choice"] in ("lane", "team") for v in rep.values())
    J0, K0 = plan.jk(D)
    sq0, pq0 = plan.last_build_stats()
    plan.set_mixed_precision(1e-6)
    J1, K1 = plan.jk(D)
    st = plan.last_mixed_stats()
    sq1, pq1 = plan.last_build_stats()
    plan.set_mixed_precision(0.0)
    ej, ek = np.abs(J1 - J0).max(), np.abs(K1 - K0).max()
    print(f"water10 autotuned: FP32 share {st['fp32_shell_quartets'] / sq1:.3f}; max|dJ| {ej:.2e} max|dK| {ek:.2e}")
    assert (sq1, pq1) == (sq0, pq0)
    assert 0 <= st["fp32_shell_quartets"] <= sq1
    assert ej <= 1e-7 and ek <= 1e-7


def test_autotune_error_leaves_plan_unchanged(monkeypatch):
    """An autotuner that fails part-way (forced from the host after it has
    switched every class to the team evaluator and the screens off) restores
    the plan: the next build evaluates the same quartets with the same
    screening and precision settings and gives the same J/K."""
    from tlno_amd import learned
    from tlno_amd._lib import TlnoGpuError
    params, margin = learned.load_weights()
    b, D = configs.water_cluster(10)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    plan.set_mixed_precision(1e-6)
    plan.set_learned_screen(params, margin)
    J0, K0 = plan.jk(D, density_threshold=2e-9)
    st0, r0, m0 = plan.last_build_stats(), plan.last_route_stats(), plan.last_mixed_stats()
    monkeypatch.setenv("TLNO_TEST_AUTOTUNE_FAIL", "1")
    with pytest.raises(TlnoGpuError, match="autotune interrupted"):
        plan.autotune(schwarz=1e-12, repeats=1)
    monkeypatch.delenv("TLNO_TEST_AUTOTUNE_FAIL")
    J1, K1 = plan.jk(D, density_threshold=2e-9)
    assert plan.last_build_stats() == st0 and plan.last_route_stats() == r0
    assert plan.last_mixed_stats()["fp32_shell_quartets"] == m0["fp32_shell_quartets"]
    assert np.abs(J1 - J0).max() < 1e-13 and np.abs(K1 - K0).max() < 1e-13


def test_persistent_launch_path_matches(monkeypatch):
    """Class launches above the grid cap run persistent (a capped grid
    striding over the work items); forced here with a cap of 64 waves, for
    the lane and the team evaluators, plain and routed builds."""
    b, D = configs.water_cluster(10)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    for team in (0, 30):
        plan.set_team_threshold(team)
        ref = [plan.jk(D), plan.jk(D, density_threshold=2e-9)]
        counts = plan.last_build_stats()
        monkeypatch.setenv("TLNO_MAX_GRID_WAVES", "64")
        got = [plan.jk(D), plan.jk(D, density_threshold=2e-9)]
        assert plan.last_build_stats() == counts
        monkeypatch.delenv("TLNO_MAX_GRID_WAVES")
        for (Ja, Ka), (Jb, Kb) in zip(ref, got):
            assert np.abs(Ja - Jb).max() < 1e-12 and np.abs(Ka - Kb).max() < 1e-12
