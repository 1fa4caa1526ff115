"""GPU against the reference's own outputs (golden fixtures written from the
reference compiled from its sources), through the C ABI.

  water7      J/K/G at Schwarz 0 and 1e-12 (spherical = Cartesian for s/p),
              S/T/V, every canonical ERI of the reference's stream
  organic5    spherical J/K/G with Cartesian-generated d shells
  organic30   the benchmark molecule in the reference's spherical basis:
              G in full and bilinear probes of J and K, Schwarz 1e-12
  density     density-screened J/K/G of a 4-water cluster at three
              density thresholds
Tolerance for J/K/G: max-abs 1e-10.
"""
import os

import numpy as np
import pytest

from tlno_amd import configs, onee
from tlno_amd.fock import DirectPlan, ScreeningOptions, build_rhf_g

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


@pytest.mark.parametrize("thr,tag", [(0.0, "0"), (1e-12, "12")])
def test_water7_jk(thr, tag):
    g = load("water7.npz")
    b, D = configs.water7()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=thr))
    J, K = plan.jk(D)
    assert np.abs(J - g[f"J_{tag}"]).max() < 1e-10
    assert np.abs(K - g[f"K_{tag}"]).max() < 1e-10
    assert np.abs(build_rhf_g(plan, D).g - g[f"G_{tag}"]).max() < 1e-10


def test_water7_eris_vs_reference_stream():
    g = load("water7.npz")
    b, _ = configs.water7()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0), cartesian=True)
    ng = plan.profile.shell_groups
    comps = plan.group_components()
    # Cartesian = spherical for s/p; assemble the full tensor from GPU blocks
    M = b.n
    E = np.zeros((M,) * 4)
    quads = [(a, bb, c, d) for a in range(ng) for bb in range(ng) for c in range(ng) for d in range(ng)]
    for q, blk in zip(quads, plan.eri_cart_blocks(quads)):
        idx = [sorted(comps[x]) for x in q]
        E[np.ix_(*idx)] = blk

    def pair(i):
        hi = int((np.sqrt(8 * i + 1) - 1) // 2)
        while hi * (hi + 1) // 2 > i:
            hi -= 1
        while (hi + 1) * (hi + 2) // 2 <= i:
            hi += 1
        return hi, i - hi * (hi + 1) // 2

    vals = np.array([E[pair(l) + pair(r)] for l, r in zip(g["stream_lhs"], g["stream_rhs"])])
    ref = g["stream_val"]
    assert np.all(np.abs(vals - ref) <= 1e-14 + 1e-12 * np.abs(ref))
    assert vals.sum() == pytest.approx(float(g["checksum"]), rel=1e-12)


def test_water7_onee():
    g = load("water7.npz")
    b, _ = configs.water7()
    for sph in (0, 1):
        plan = DirectPlan(b, ScreeningOptions(), cartesian=not sph)
        m = onee.evaluate(plan, b.nuclear_charges, b.nuclear_centers)
        for x, r in ((m.overlap, g[f"S{sph}"]), (m.kinetic, g[f"T{sph}"]), (m.nuclear, g[f"V{sph}"])):
            assert np.all(np.abs(x - r) <= 1e-14 + 1e-12 * np.abs(r) + 1e-14 * np.abs(r).max())


def test_organic5_spherical_jk():
    g = load("organic5.npz")
    b, _ = configs.organic30(natoms=5)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=0.0))
    J, K = plan.jk(g["Dsph"])
    assert np.abs(J - g["J_0"]).max() < 1e-10
    assert np.abs(K - g["K_0"]).max() < 1e-10
    assert np.abs(J - 0.5 * K - g["G_0"]).max() < 1e-10


def test_organic30_benchmark_molecule_vs_reference():
    g = load("organic30.npz")
    b, _ = configs.organic30()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12))
    D = configs.spherical_density(plan.M, 5, 60)
    J, K = plan.jk(D)
    assert np.abs(J - 0.5 * K - g["G_12"]).max() < 1e-10
    U, V = g["U"], g["Vp"]
    jp = np.einsum("im,mn,in->i", U, J, V)
    kp = np.einsum("im,mn,in->i", U, K, V)
    assert np.allclose(jp, g["J_probe"], rtol=0, atol=1e-9)
    assert np.allclose(kp, g["K_probe"], rtol=0, atol=1e-9)
    assert np.linalg.norm(J) == pytest.approx(float(g["J_fro"]), rel=1e-12)
    assert np.linalg.norm(K) == pytest.approx(float(g["K_fro"]), rel=1e-12)


def test_organic30_autotuned_vs_reference():
    """The benchmark molecule with the autotuned plan (team evaluator on most
    d class pairs, quartet lists, cost-balanced lane slices) against the
    reference's G at Schwarz 1e-12."""
    g = load("organic30.npz")
    b, _ = configs.organic30()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12))
    rep = plan.autotune(schwarz=1e-12, repeats=1)
    assert any(v["choice"] == "team" for v in rep.values())
    D = configs.spherical_density(plan.M, 5, 60)
    J, K = plan.jk(D)
    assert np.abs(J - 0.5 * K - g["G_12"]).max() < 1e-10
    assert np.linalg.norm(J) == pytest.approx(float(g["J_fro"]), rel=1e-12)
    assert np.linalg.norm(K) == pytest.approx(float(g["K_fro"]), rel=1e-12)


@pytest.mark.parametrize("tau,tag", [(1e-9, "9"), (1e-7, "7"), (1e-5, "5")])
def test_density_screen_vs_reference(tau, tag):
    """The reference's density-screened J (J-only test), K (K-only test) and
    G (combined test): the per-quartet test keeps the same quartets, so the
    results agree to 1e-10 even where the screen changes G by 1e-5."""
    g = load("density.npz")
    b, _ = configs.water_cluster(4)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12))
    Ds = configs.spherical_density(plan.M, 9, 20)
    J, _ = plan.jk(Ds, True, False, density_threshold=tau)
    _, K = plan.jk(Ds, False, True, density_threshold=tau)
    G = build_rhf_g(plan, Ds, ScreeningOptions(integral_threshold=1e-12, density_threshold=tau)).g
    print(f"tau {tau:g}: screen moves G by {np.abs(g[f'G_{tag}'] - g['G_0']).max():.2e}; "
          f"|dJ| {np.abs(J - g[f'J_{tag}']).max():.2e} |dK| {np.abs(K - g[f'K_{tag}']).max():.2e} "
          f"|dG| {np.abs(G - g[f'G_{tag}']).max():.2e}")
    assert np.abs(J - g[f"J_{tag}"]).max() < 1e-10
    assert np.abs(K - g[f"K_{tag}"]).max() < 1e-10
    assert np.abs(G - g[f"G_{tag}"]).max() < 1e-10


def test_eri_stream_checksums_vs_reference():
    """The reference's packed-stream checksum (sum of the unique spherical
    ERIs): water7 and organic5 unscreened, organic30 at Schwarz 1e-12 and
    1e-8.  Relative tolerance 1e-11 (a sum of ~1e8 terms in another order)."""
    from tlno_amd.fock import build_eri_packed_for_G_rys, EriPackedBuildOptions, build_eri_packed_plan_for_G_rys
    g7, g5, gc = load("water7.npz"), load("organic5.npz"), load("checksum.npz")
    b7, _ = configs.water7()
    p7 = build_eri_packed_plan_for_G_rys(b7, EriPackedBuildOptions(schwarz_threshold=0.0))
    c7 = build_eri_packed_for_G_rys(p7, EriPackedBuildOptions(schwarz_threshold=0.0)).values[0]
    b5, _ = configs.organic30(natoms=5)
    p5 = DirectPlan(b5, ScreeningOptions(integral_threshold=0.0))
    c5 = p5.eri_checksum(0.0)
    b30, _ = configs.organic30()
    p30 = DirectPlan(b30, ScreeningOptions(integral_threshold=1e-12))
    c30 = {t: p30.eri_checksum(thr) for t, thr in (("12", 1e-12), ("8", 1e-8))}
    print(f"water7 {c7!r} vs {float(g7['checksum'])!r}; organic5 {c5!r} vs {float(g5['checksum'])!r}; "
          f"organic30 {c30} vs {float(gc['organic30_12'])!r} {float(gc['organic30_8'])!r}")
    assert c7 == pytest.approx(float(g7["checksum"]), rel=1e-11)
    assert c5 == pytest.approx(float(g5["checksum"]), rel=1e-11)
    assert c30["12"] == pytest.approx(float(gc["organic30_12"]), rel=1e-11)
    assert c30["8"] == pytest.approx(float(gc["organic30_8"]), rel=1e-11)
    with pytest.raises(Exception, match="pure spherical"):
        DirectPlan(b7, cartesian=True).eri_checksum(0.0)


def test_eri_stream_fingerprints_vs_reference():
    """The reference's indexed stream fingerprint (its
    TLNO_VALIDATE_RYS_STREAM_FINGERPRINT build: checksum, three moments of
    the values weighted by hashes of their canonical quartet index, count) on
    water7 and organic5 unscreened, organic5 and organic30 at Schwarz 1e-12.
    The count is exact; the sums agree to 1e-12 of sum |x| (a value in the
    wrong slot moves a moment by about its own size).  water7's unindexed
    components are checked against the restated fingerprint of the
    reference's recorded stream, to the bounds that the per-value tolerance
    1e-14 + 1e-12 |x| implies."""
    from oracle import oracle as O
    from tlno_amd.fock import build_eri_packed_for_G_rys, EriPackedBuildOptions, build_eri_packed_plan_for_G_rys
    f = load("fingerprint.npz")
    cases = [("water7", configs.water7()[0], 0.0, "0"), ("organic5", configs.organic30(natoms=5)[0], 0.0, "0"),
             ("organic5", configs.organic30(natoms=5)[0], 1e-12, "12"), ("organic30", configs.organic30()[0], 1e-12, "12")]
    for name, b, thr, tag in cases:
        plan = DirectPlan(b, ScreeningOptions(integral_threshold=thr))
        fp = plan.eri_fingerprint(thr)
        ref = f[f"{name}_{tag}"]
        tol = 1e-12 * fp["absolute_sum"]
        d = [abs(fp[k] - ref[i]) for i, k in enumerate(("checksum", "indexed_moment_1", "indexed_moment_2",
                                                          "indexed_moment_3"))]
        print(f"{name} at {thr}: {fp['canonical_count']} values (reference {int(ref[4])}); "
              f"max |d| {max(d):.2e} of sum|x| {fp['absolute_sum']:.4e}")
        assert fp["canonical_count"] == int(ref[4])
        assert max(d) <= tol
        assert fp["checksum"] == pytest.approx(plan.eri_checksum(thr), rel=1e-12)
    g7 = load("water7.npz")
    o7 = O.stream_fingerprint(g7["stream_lhs"], g7["stream_rhs"], g7["stream_val"])
    b7, _ = configs.water7()
    fp7 = DirectPlan(b7, ScreeningOptions(integral_threshold=0.0)).eri_fingerprint(0.0)
    # per value |d| <= 1e-14 + 1e-12 |x| bounds the sums of |x| and x^2 and the maximum
    n7 = o7["canonical_count"]
    assert abs(fp7["absolute_sum"] - o7["absolute_sum"]) <= n7 * 1e-14 + 1e-12 * o7["absolute_sum"]
    assert abs(fp7["square_sum"] - o7["square_sum"]) <= 2 * o7["max_abs"] * (n7 * 1e-14 + 1e-12 * o7["absolute_sum"])
    assert abs(fp7["max_abs"] - o7["max_abs"]) <= 1e-14 + 1e-12 * o7["max_abs"]
    # the packed endpoint's fingerprint form returns what the reference's build returns
    p7 = build_eri_packed_plan_for_G_rys(b7, EriPackedBuildOptions(schwarz_threshold=0.0))
    v = build_eri_packed_for_G_rys(p7, EriPackedBuildOptions(schwarz_threshold=0.0), fingerprint=True).values
    assert len(v) == 5 and v[4] == 406.0
    assert np.abs(np.array(v[:4]) - f["water7_0"][:4]).max() <= 1e-12 * fp7["absolute_sum"]
    with pytest.raises(Exception, match="pure spherical"):
        DirectPlan(b7, cartesian=True).eri_fingerprint(0.0)
