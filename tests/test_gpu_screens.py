"""GPU: the reference's heuristic screens in the J/K builds (ScreeningOptions
decay_threshold and primitive_fastschwarz_threshold), against the reference's
own screened builds (tests/golden/screens.npz) and the oracle's restatement
(oracle/orc_eri.c orc_jk_screened, itself pinned to the reference in
tests/test_oracle_screens.py):

  * decay indicator, with Schwarz 1e-12 and alone: the kept shell quartets
    are the reference's, G within 1e-10 of the reference's;
  * primitive FASTSCHWARZ (per primitive quartet here): G within 1e-10 of the
    reference on configs.distinct_pairs() (one pair per SIMD packet of the
    reference, where the two rules coincide), and J/K within 1e-10 of the
    oracle's per-quartet rule on 4 waters, lane and team evaluators.  The
    reference's packet rule keeps more primitive quartets; the difference is
    printed.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tlno_amd import configs
from tlno_amd.fock import DirectPlan, ScreeningOptions, build_rhf_g

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "screens.npz"))


@pytest.mark.parametrize("sch,dec,tag", [(1e-12, 1e-6, "s12_d6"), (1e-12, 1e-4, "s12_d4"), (0.0, 1e-8, "d8"),
                                         (0.0, 1e-5, "d5")])
def test_decay_screen_matches_reference(sch, dec, tag):
    b, _ = configs.water_cluster(4)
    opts = ScreeningOptions(integral_threshold=sch, decay_threshold=dec)
    plan = DirectPlan(b, opts)
    D = configs.spherical_density(plan.M, 9, 20)
    G = build_rhf_g(plan, D).g
    ref = GOLD[f"water4_G_{tag}"]
    d = float(np.abs(G - ref).max())
    G0 = build_rhf_g(plan, D, ScreeningOptions(integral_threshold=sch)).g
    print(f"decay {dec:g} schwarz {sch:g}: |G - G_ref| {d:.2e}; the screen moves G by {np.abs(G0 - ref).max():.2e}")
    assert d <= 1e-10


def test_decay_screen_cartesian_matches_oracle():
    b, D = configs.water_cluster(4)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J, K = plan.jk(D, decay=1e-5)
    sq_dec, _ = plan.last_build_stats()
    J0, K0 = plan.jk(D)
    sq0, _ = plan.last_build_stats()
    Jr, Kr = O.OracleBasis(b).jk(D, schwarz=1e-12, decay=1e-5)
    assert sq_dec < sq0
    assert np.abs(J - Jr).max() <= 1e-10 and np.abs(K - Kr).max() <= 1e-10
    assert max(np.abs(J - J0).max(), np.abs(K - K0).max()) > 1e-9  # the screen is active


def test_fastschwarz_matches_reference_with_single_pair_packets():
    b, D = configs.distinct_pairs()
    for team in (0, 1):
        plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0))
        plan.set_team_threshold(team)
        for tag, fs in (("6", 1e-6), ("4", 1e-4), ("2", 1e-2)):
            G = build_rhf_g(plan, D, ScreeningOptions(integral_threshold=0.0, primitive_fastschwarz_threshold=fs)).g
            assert np.abs(G - GOLD[f"distinct_G_fs{tag}"]).max() <= 1e-10, (team, tag)
        G0 = build_rhf_g(plan, D, ScreeningOptions(integral_threshold=0.0)).g
        assert np.abs(G0 - GOLD["distinct_G_0"]).max() <= 1e-10


@pytest.mark.parametrize("team", [0, 30])
@pytest.mark.parametrize("fs", [1e-9, 1e-7])
def test_fastschwarz_matches_oracle_rule(fs, team):
    b, D = configs.water_cluster(4)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    plan.set_team_threshold(team)
    J, K = plan.jk(D, fastschwarz=fs)
    Jr, Kr = O.OracleBasis(b).jk(D, schwarz=1e-12, fastschwarz=fs)
    J0, K0 = plan.jk(D)
    e = max(np.abs(J - Jr).max(), np.abs(K - Kr).max())
    moved = max(np.abs(J - J0).max(), np.abs(K - K0).max())
    print(f"fastschwarz {fs:g} team>={team}: |JK - oracle| {e:.2e}; the screen moves J/K by {moved:.2e}")
    assert e <= 1e-10
    assert moved > 1e-9


def test_fastschwarz_against_reference_packet_rule():
    """4 waters, spherical: the reference keeps every primitive quartet of a
    SIMD packet of bra pairs when one pair of the packet passes; this engine
    decides per quartet.  Both stay close to the unscreened G; this engine's
    screen moves G by no more than about ten times the reference's."""
    b, _ = configs.water_cluster(4)
    plan = DirectPlan(b)
    D = configs.spherical_density(plan.M, 9, 20)
    G0 = build_rhf_g(plan, D).g
    for tag, fs in (("9", 1e-9), ("7", 1e-7)):
        G = build_rhf_g(plan, D, ScreeningOptions(primitive_fastschwarz_threshold=fs)).g
        ref = GOLD[f"water4_G_fs{tag}"]
        ours, theirs = float(np.abs(G - G0).max()), float(np.abs(ref - G0).max())
        print(f"fastschwarz {fs:g}: |G - G_unscreened| {ours:.2e} (reference {theirs:.2e}), |G - G_ref| "
              f"{np.abs(G - ref).max():.2e}")
        assert theirs <= ours <= 20 * theirs + 1e-12


def test_screen_option_errors():
    b, _ = configs.water_cluster(2)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    D = np.eye(plan.M) * 1e-2
    with pytest.raises(RuntimeError, match="cannot be combined"):
        plan.jk(D, density_threshold=1e-9, decay=1e-8)
    with pytest.raises(RuntimeError, match="decay_threshold must be finite"):
        plan.jk(D, decay=-1.0)
    with pytest.raises(RuntimeError, match="primitive_fastschwarz_threshold must be finite"):
        plan.jk(D, fastschwarz=float("nan"))
    # settings are per build: a following plain build is unscreened
    J1, K1 = plan.jk(D, fastschwarz=1e-6, decay=1e-6)
    J0, K0 = plan.jk(D)
    J2, K2 = plan.jk(D, fastschwarz=0.0, decay=0.0)
    assert np.abs(J2 - J0).max() < 1e-13 and np.abs(K2 - K0).max() < 1e-13
