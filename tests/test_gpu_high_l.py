"""GPU parity for f and g shells (L = 3, 4): the team evaluator of every class
pair with an f or a g shell, the Rys rules of 6..9 roots, the f and g
one-electron kernels and the spherical f and g projections, against the
oracle, with the reference's own outputs beside them (tests/golden/fshell.npz:
configs.organic_f(), 125 Cartesian / 105 spherical functions; gshell.npz:
configs.organic_g(), 120 / 90; written by tests/golden/gen_golden.py from the
reference built from its sources and from the oracle).  Every test but the Rys
one runs once per molecule.

Tolerances:
  Rys rules, 6..9 roots   relative 1e-13 against the oracle.  The oracle's
        rules reproduce 40-digit Boys moments to 1e-15 (tests/test_oracle_high_l.py);
        the reference's table route is good to about 1e-12 in the moments and
        2e-11..9e-11 in the roots for N = 7..9, which the tests print.
  ERI blocks   |d| <= 1e-14 + 1e-12 |ref| against the oracle, an entry outside
        that bound judged against its 40-digit value (as test_gpu_parity);
        for g blocks the bound adds 1e-10 of the block's largest entry (the
        GPU's worst is 4e-11 of it, the reference's audit route's 6e-10).
  J/K    max-abs <= 1e-10 against the oracle (Cartesian and spherical) and
        against the reference's spherical J, K and G.
  density screen   the change the screen makes to G, GPU against reference,
        max-abs <= 1e-10 (the reference's Rys error cancels in the difference,
        so the kept quartet sets must be identical).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tlno_amd import configs, core_math, onee
from tlno_amd.fock import DirectPlan, ScreeningOptions

from test_gpu_parity import ExactFallback, _block_exact, _close
from test_oracle_high_l import oriented

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
JK_ABS = 1e-10


CASES = {"f": ("fshell.npz", configs.organic_f), "g": ("gshell.npz", configs.organic_g)}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    return request.param


@pytest.fixture(scope="module")
def fx(case):
    return np.load(os.path.join(GOLD, CASES[case][0]))


@pytest.fixture(scope="module")
def mol(case, fx):
    b, D = CASES[case][1]()
    from test_oracle import basis_hash
    assert basis_hash(b) == str(fx["basis_sha"])
    return b, D


@pytest.mark.parametrize("N", [6, 7, 8, 9])
def test_rys_high_orders_match_oracle(N):
    fx = np.load(os.path.join(GOLD, "fshell.npz"))
    rng = np.random.default_rng(100 + N)
    edges = [0.0, 1e-9, 0.5, 1.0, 29.0, 49.0, 66.99, 67.0, 71.99, 72.0, 76.99, 77.0, 82.99, 83.0, 87.99, 88.0,
             120.0, 5000.0]
    T = np.concatenate([edges, rng.uniform(0, 95, 300), fx["Ts"]])
    r, w = core_math.rys(N, T)
    for i, t in enumerate(T):
        rr, ww = O.rys(N, t)
        assert np.allclose(r[i], rr, rtol=1e-13, atol=0), (N, t, r[i], rr)
        assert np.allclose(w[i], ww, rtol=1e-13, atol=0), (N, t, w[i], ww)
    n = len(fx["Ts"])
    dr = np.abs(r[-n:] - fx[f"rys_r{N}"]) / np.abs(fx[f"rys_r{N}"])
    dw = np.abs(w[-n:] - fx[f"rys_w{N}"]) / np.abs(fx[f"rys_w{N}"])
    print(f"N={N}: GPU against the reference's table route: roots {dr.max():.2e}, weights {dw.max():.2e}")
    assert max(dr.max(), dw.max()) < 1e-9


# g blocks: the bra and ket transfers (HRR) of up to eight units of angular
# momentum cancel digits in the small entries of a block, up to about 4e-11 of
# its largest entry (the reference's audit route loses up to 6e-10 on the same
# blocks; both are printed).  The bound adds a floor of 1e-10 of the block's
# largest entry for g blocks only.
HRR_FLOOR = {"f": 0.0, "g": 1e-10}


def test_blocks_match_oracle_and_reference(case, fx, mol):
    b, _ = mol
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0), cartesian=True)
    ob = O.OracleBasis(b)
    quads = [tuple(int(x) for x in q) for q in fx["quads"]]
    blocks = plan.eri_cart_blocks(quads)
    fb = ExactFallback()
    outside_ref = 0
    worst_gpu = worst_ref = 0.0
    for k, (q, blk) in enumerate(zip(quads, blocks)):
        o = oriented(ob, q)
        ref = fx["blocks"][fx["block_off"][k]:fx["block_off"][k + 1]].reshape(blk.shape)
        scale = np.abs(o).max()
        worst_gpu = max(worst_gpu, np.abs(blk - o).max() / scale)
        worst_ref = max(worst_ref, np.abs(ref - o).max() / scale)
        fb.check(blk, o, _block_exact(b, ob, q), floor=HRR_FLOOR[case] * scale)
        # the reference's audit route: same convention and order (an indexing
        # mistake would show at the size of the entries)
        assert np.abs(blk - ref).max() <= 1e-8 * max(1.0, np.abs(ref).max()), q
        outside_ref += int(np.count_nonzero(np.abs(blk - ref) > 1e-14 + 1e-12 * np.abs(ref)))
    print(f"{len(quads)} blocks ({min(x.size for x in blocks)}-{max(x.size for x in blocks)} values); "
          f"{fb.used} entries judged against their 40-digit values; {outside_ref} entries outside the strict "
          f"bound of the reference's audit blocks; largest error over the block's largest entry: "
          f"GPU {worst_gpu:.2e}, reference {worst_ref:.2e}")


def test_jk_cartesian_matches_oracle(fx, mol):
    b, D = mol
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J, K = plan.jk(D, schwarz=1e-12)
    ej = np.abs(J - fx["Jc_12"]).max()
    ek = np.abs(K - fx["Kc_12"]).max()
    print(f"Cartesian J/K against the oracle: {ej:.2e} {ek:.2e}")
    assert ej < JK_ABS and ek < JK_ABS


def test_jk_spherical_matches_oracle_and_reference(fx, mol):
    b, _ = mol
    Ds = fx["Dsph"]
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=False)
    J, K = plan.jk(Ds, schwarz=1e-12)
    ej = np.abs(J - fx["Jo_12"]).max()
    ek = np.abs(K - fx["Ko_12"]).max()
    G = J - 0.5 * K
    eg = np.abs(G - fx["G_12"]).max()
    ejr = np.abs(J - fx["J_12"]).max()
    ekr = np.abs(K - fx["K_12"]).max()
    go = fx["Jo_12"] - 0.5 * fx["Ko_12"]
    print(f"spherical J/K against the oracle: {ej:.2e} {ek:.2e}; against the reference J {ejr:.2e}, "
          f"K {ekr:.2e}, G {eg:.2e} (oracle against the reference: G {np.abs(go - fx['G_12']).max():.2e})")
    assert ej < JK_ABS and ek < JK_ABS
    assert max(eg, ejr, ekr) < JK_ABS


def test_density_screen_matches_reference(fx, mol):
    b, _ = mol
    Ds = fx["Dsph"]
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=False)
    J0, K0 = plan.jk(Ds, schwarz=1e-12)
    n0 = plan.last_build_stats()[0]
    J1, K1 = plan.jk(Ds, schwarz=1e-12, density_threshold=1e-7)
    n1 = plan.last_build_stats()[0]
    dg = (J1 - 0.5 * K1) - (J0 - 0.5 * K0)
    dref = fx["G_d7"] - fx["G_12"]
    err = np.abs(dg - dref).max()
    print(f"density screen 1e-7: {n0} -> {n1} shell quartets; G moved by {np.abs(dref).max():.2e}, "
          f"GPU against reference change {err:.2e}")
    assert n1 < n0
    assert np.abs(dref).max() > 1e-9
    assert err < JK_ABS


def test_heuristic_screens_match_oracle(case):
    b, D = CASES[case][1](natoms=3)
    ob = O.OracleBasis(b)
    for kw in ({"decay": 1e-6}, {"fastschwarz": 1e-7}):
        plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12,
                                              decay_threshold=kw.get("decay", 0.0),
                                              primitive_fastschwarz_threshold=kw.get("fastschwarz", 0.0)),
                          cartesian=True)
        J, K = plan.jk(D, schwarz=1e-12)
        Jr, Kr = ob.jk(D, spherical=False, schwarz=1e-12, **kw)
        J0, K0 = ob.jk(D, spherical=False, schwarz=1e-12)
        print(f"{kw}: screen moves J/K by {np.abs(Jr - J0).max():.2e} {np.abs(Kr - K0).max():.2e}; "
              f"GPU against oracle {np.abs(J - Jr).max():.2e} {np.abs(K - Kr).max():.2e}")
        assert np.abs(J - Jr).max() < JK_ABS and np.abs(K - Kr).max() < JK_ABS


def test_persistent_stride(monkeypatch, mol):
    """Capped grids: the f- and g-class team launches stride over their quartets."""
    b, D = mol
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    J0, K0 = plan.jk(D, schwarz=1e-12)
    monkeypatch.setenv("TLNO_MAX_GRID_WAVES", "37")
    J1, K1 = plan.jk(D, schwarz=1e-12)
    assert np.abs(J1 - J0).max() < 1e-12 and np.abs(K1 - K0).max() < 1e-12


@pytest.mark.parametrize("cartesian", [True, False])
def test_onee_matches_oracle(fx, mol, cartesian):
    b, _ = mol
    plan = DirectPlan(b, ScreeningOptions(), cartesian=cartesian)
    m = onee.evaluate(plan, b.nuclear_charges, b.nuclear_centers)
    ob = O.OracleBasis(b)
    S, T, V = ob.onee(b.nuclear_charges, b.nuclear_centers, spherical=not cartesian)
    import mp_integrals as mpi
    Z, R = b.nuclear_charges, b.nuclear_centers
    fns = (lambda i: mpi.overlap(b, *i), lambda i: mpi.kinetic(b, *i), lambda i: mpi.nuclear(b, *i, Z, R))
    fb = ExactFallback()
    tag = "0" if cartesian else "1"
    for name, x, r, f in zip("STV", (m.overlap, m.kinetic, m.nuclear), (S, T, V), fns):
        if cartesian:
            fb.check(x, r, f)
        else:
            d = np.sqrt(np.abs(np.diag(r)))
            ok, worst = _close(x, r, floor=1e-13 * np.outer(d, d))
            assert ok, (name, worst)
        print(f"{name} ({'Cartesian' if cartesian else 'spherical'}): against the reference "
              f"{np.abs(x - fx[name + tag]).max():.2e}")
    # S and V agree with the reference to its rounding; T only to its
    # one-sided kinetic form's cancellation, printed above
    for name, x in (("S", m.overlap), ("V", m.nuclear)):
        r = fx[name + tag]
        assert np.abs(x - r).max() < 1e-9 * max(1.0, np.abs(r).max()), name


def test_checksum_matches_reference(fx, mol):
    b, _ = mol
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=False)
    c = plan.eri_checksum(1e-12)
    ref = float(fx["checksum_12"])
    print(f"stream checksum {c!r} against the reference's {ref!r}: relative {abs(c - ref) / abs(ref):.2e}")
    assert abs(c - ref) <= 1e-10 * abs(ref)
