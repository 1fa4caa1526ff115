"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs.

Tolerances (stated per test):
  ERIs, S/T/V   |d| <= 1e-14 + 1e-12 |ref| against the oracle; an entry
         outside that bound (at most a few, entries that vanish or cancel)
         must meet the same bound against its 40-digit value
         (tests/mp_integrals.py, McMurchie-Davidson in mpmath).  Spherical
         S/T/V (combinations of Cartesian entries) add the combination's
         rounding, 1e-13 sqrt|X_ss X_tt|.
  J/K    max-abs <= 1e-10.
  Boys and Rys rules: relative 1e-13.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tlno_amd import configs, core_math
from tlno_amd.fock import DirectPlan, ScreeningOptions, build_coulomb_j, build_exchange_k, build_rhf_g
from tlno_amd import onee

pytestmark = pytest.mark.gpu

ERI_ABS, ERI_REL = 1e-14, 1e-12
JK_ABS = 1e-10


def _close(x, ref, abs_tol=ERI_ABS, rel_tol=ERI_REL, floor=0.0):
    x = np.asarray(x)
    ref = np.asarray(ref)
    tol = abs_tol + rel_tol * np.abs(ref) + floor
    bad = np.abs(x - ref) > tol
    return not bad.any(), float(np.max(np.abs(x - ref) / tol))


class ExactFallback:
    """Strict comparison with the oracle; entries outside the bound are
    judged against their 40-digit values instead (at most `limit`)."""

    def __init__(self, limit=24):
        self.limit = limit
        self.used = 0

    def check(self, x, ref, exact_of, floor=0.0):
        x = np.asarray(x)
        ref = np.asarray(ref)
        bad = np.argwhere(np.abs(x - ref) > ERI_ABS + ERI_REL * np.abs(ref) + floor)
        for idx in bad:
            idx = tuple(int(i) for i in idx)
            self.used += 1
            assert self.used <= self.limit, "too many entries differ from the oracle"
            ex = float(exact_of(idx))
            assert abs(x[idx] - ex) <= ERI_ABS + ERI_REL * abs(ex) + floor, (idx, x[idx], ref[idx], ex)


def _block_exact(b, ob, q):
    import mp_integrals as mpi
    comps = [ob.groups[g][1] for g in q]
    return lambda idx: mpi.eri(b, *[int(comps[m][idx[m]]) for m in range(4)])


@pytest.fixture(scope="module")
def small_organic():
    return configs.organic30(natoms=5)


def test_boys_matches_oracle():
    rng = np.random.default_rng(1)
    T = np.concatenate([[0.0, 1e-12, 1e-6, 0.0625, 0.125, 39.999, 40.0, 40.001, 55.0, 120.0, 1e4],
                        rng.uniform(0, 45, 400), rng.uniform(0, 2, 100)])
    for mmax in (0, 1, 4, 8, 16):
        F = core_math.boys(T, mmax)
        ref = np.array([[O.boys(m, t) for m in range(mmax + 1)] for t in T])
        rel = np.abs(F - ref) / np.abs(ref)
        assert rel.max() < 1e-13, (mmax, rel.max(), T[np.unravel_index(rel.argmax(), rel.shape)[0]])


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_rys_matches_oracle(N):
    rng = np.random.default_rng(N)
    edges = [0.0, 1e-9, 0.5, 0.999999, 1.0, 1.000001, 12.5, 29.0, 37.0, 41.0, 48.99, 49.0, 54.99, 55.0,
             60.99, 61.0, 66.99, 67.0, 80.0, 200.0, 5000.0]
    T = np.concatenate([edges, rng.uniform(0, 70, 300)])
    r, w = core_math.rys(N, T)
    for i, t in enumerate(T):
        rr, ww = O.rys(N, t)
        assert np.allclose(r[i], rr, rtol=1e-13, atol=0), (N, t, r[i], rr)
        assert np.allclose(w[i], ww, rtol=1e-13, atol=0), (N, t, w[i], ww)


def _boys_exact(k, T):
    import mpmath as mp
    T = mp.mpf(T)
    if T == 0:
        return mp.mpf(1) / (2 * k + 1)
    return mp.gammainc(k + mp.mpf(1) / 2, 0, T) / (2 * T ** (k + mp.mpf(1) / 2))


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_rys_gpu_moments_exact(N):
    """The GPU rule against the mathematics, not the oracle: sum_i w_i t_i^k =
    F_k(T) for k < 2N with F_k from the incomplete gamma function at 30
    digits, independent of the start-up tables and of the oracle.  The rules
    measure within 2.5e-16 (k + 2) of the moments; the bound allows 2e-15
    (k + 2)."""
    import mpmath as mp
    mp.mp.dps = 30
    T = np.concatenate([[0.0, 1e-9, 0.999999, 1.0, 29.0, 41.0, 49.0, 55.0, 61.0, 67.0, 77.0, 88.0, 200.0, 5000.0],
                        np.random.default_rng(300 + N).uniform(0, 95, 24)])
    r, w = core_math.rys(N, T)
    worst = 0.0
    for i, t in enumerate(T):
        assert np.all(np.diff(r[i]) > 0) and r[i][0] >= 0 and r[i][-1] < 1 and np.all(w[i] > 0), (N, t)
        for k in range(2 * N):
            ex = _boys_exact(k, t)
            mom = mp.fsum(mp.mpf(float(w[i][j])) * mp.mpf(float(r[i][j])) ** k for j in range(N))
            rel = float(abs(mom - ex) / ex)
            worst = max(worst, rel / (k + 2))
            assert rel <= (k + 2) * 2e-15, (N, t, k, rel)
    print(f"N={N}: worst moment error / (k + 2) {worst:.2e}")


def test_eri_blocks_water7_all():
    b, _ = configs.water7()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0), cartesian=True)
    ob = O.OracleBasis(b)
    ng = ob.ngroups
    quads = [(a, bb, c, d) for a in range(ng) for bb in range(ng) for c in range(ng) for d in range(ng)]
    blocks = plan.eri_cart_blocks(quads)
    fb = ExactFallback()
    for q, blk in zip(quads, blocks):
        fb.check(blk, ob.cart_block(*q), _block_exact(b, ob, q))


def test_eri_blocks_with_d_shells(small_organic):
    b, _ = small_organic
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0), cartesian=True)
    ob = O.OracleBasis(b)
    Ls = plan.group_L()
    ng = len(Ls)
    rng = np.random.default_rng(7)
    quads = set()
    # every class combination at least once, plus random quadruples
    byL = {L: [g for g in range(ng) if Ls[g] == L] for L in set(Ls)}
    for la in byL:
        for lb in byL:
            for lc in byL:
                for ld in byL:
                    quads.add((rng.choice(byL[la]), rng.choice(byL[lb]), rng.choice(byL[lc]), rng.choice(byL[ld])))
    for _ in range(200):
        quads.add(tuple(int(x) for x in rng.integers(0, ng, 4)))
    quads = sorted(quads)
    blocks = plan.eri_cart_blocks(quads)
    fb = ExactFallback()
    for q, blk in zip(quads, blocks):
        fb.check(blk, ob.cart_block(*q), _block_exact(b, ob, q))
    print(f"{len(quads)} blocks; {fb.used} entries judged against their 40-digit values")


@pytest.mark.parametrize("cartesian", [True, False])
def test_jk_water7(cartesian):
    b, D = configs.water7()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0), cartesian=cartesian)
    ob = O.OracleBasis(b)
    Jr, Kr = ob.jk(D, spherical=not cartesian)
    J = build_coulomb_j(plan, D).j
    K = build_exchange_k(plan, D).k
    G = build_rhf_g(plan, D).g
    assert np.abs(J - Jr).max() < JK_ABS
    assert np.abs(K - Kr).max() < JK_ABS
    assert np.abs(G - (Jr - 0.5 * Kr)).max() < JK_ABS


@pytest.mark.parametrize("cartesian", [True, False])
def test_jk_small_organic(small_organic, cartesian):
    b, Dc = small_organic
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0), cartesian=cartesian)
    ob = O.OracleBasis(b)
    if cartesian:
        D = Dc
    else:
        rng = np.random.default_rng(3)
        C = rng.standard_normal((ob.nsph, 10))
        D = C @ C.T
        D /= np.linalg.norm(D)
    Jr, Kr = ob.jk(D, spherical=not cartesian)
    J, K = plan.jk(D)
    assert np.abs(J - Jr).max() < JK_ABS, np.abs(J - Jr).max()
    assert np.abs(K - Kr).max() < JK_ABS, np.abs(K - Kr).max()


def test_jk_schwarz_screened_close_to_unscreened(small_organic):
    b, D = small_organic
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J0, K0 = plan.jk(D, schwarz=0.0)
    J1, K1 = plan.jk(D, schwarz=1e-12)
    assert np.abs(J1 - J0).max() < JK_ABS
    assert np.abs(K1 - K0).max() < JK_ABS


@pytest.mark.parametrize("cartesian", [True, False])
def test_onee(small_organic, cartesian):
    b, _ = small_organic
    plan = DirectPlan(b, ScreeningOptions(), cartesian=cartesian)
    m = onee.evaluate(plan, b.nuclear_charges, b.nuclear_centers)
    ob = O.OracleBasis(b)
    S, T, V = ob.onee(b.nuclear_charges, b.nuclear_centers, spherical=not cartesian)
    import mp_integrals as mpi
    Z, R = b.nuclear_charges, b.nuclear_centers
    fns = (lambda i: mpi.overlap(b, *i), lambda i: mpi.kinetic(b, *i), lambda i: mpi.nuclear(b, *i, Z, R))
    fb = ExactFallback()
    for x, r, f in zip((m.overlap, m.kinetic, m.nuclear), (S, T, V), fns):
        if cartesian:
            fb.check(x, r, f)
        else:
            # spherical entries are combinations of up to 36 Cartesian values
            # with coefficients up to ~2, formed in a different order here and
            # in the oracle: the rounding of the combination, 1e-13 of the
            # diagonal scale sqrt|X_ss X_tt|, bounds their difference
            d = np.sqrt(np.abs(np.diag(r)))
            ok, worst = _close(x, r, floor=1e-13 * np.outer(d, d))
            assert ok, worst
    print(f"{fb.used} entries judged against their 40-digit values")


@pytest.mark.parametrize("spherical", [False, True])
def test_team_evaluator_matches_lane_evaluator(spherical):
    """The one-wave-per-quartet evaluator (HRR after contraction) against the
    one-lane-per-quartet evaluator and the oracle, every class with d shells."""
    from tlno_amd import configs
    from tlno_amd.fock import DirectPlan, ScreeningOptions
    b, D = configs.organic30(natoms=6, box_A=5.0)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0), cartesian=not spherical)
    Dx = D if not spherical else configs.spherical_density(plan.M, 4, 10)
    plan.set_team_threshold(0)
    J0, K0 = plan.jk(Dx)
    plan.set_team_threshold(1)  # every class with a team kernel
    J1, K1 = plan.jk(Dx)
    plan.set_team_threshold(216)
    assert np.abs(J1 - J0).max() < 1e-11 and np.abs(K1 - K0).max() < 1e-11
    if not spherical:
        from oracle import oracle as O
        Jo, Ko = O.OracleBasis(b).jk(D, spherical=False, schwarz=0.0)
        assert np.abs(J1 - Jo).max() < 1e-10 and np.abs(K1 - Ko).max() < 1e-10
    rep = plan.autotune(schwarz=0.0, repeats=1)
    assert rep and all(v["choice"] in ("lane", "team") for v in rep.values())
    J2, K2 = plan.jk(Dx)
    assert np.abs(J2 - J0).max() < 1e-11 and np.abs(K2 - K0).max() < 1e-11
