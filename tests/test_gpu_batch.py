"""GPU: molecule batches (persistent class launches over many molecules,
HIP-graph replay, autotuned launch shapes) against the CPU oracle per
molecule and against single-molecule plans.  Tolerances: 1e-10 max-abs
against the oracle (FP64 throughout), 1e-12 against single-molecule GPU
builds and between graph replays and eager builds."""
import numpy as np
import pytest

from oracle import oracle as O
from tlno_amd import configs, hipmem
from tlno_amd.batch import BatchPlan
from tlno_amd.fock import DirectPlan, ScreeningOptions

pytestmark = pytest.mark.gpu


def _mols(n=10, atoms=(3, 7), seed=77):
    return configs.batch256(seed=seed, nmol=n, atom_range=atoms)


def test_batch_matches_oracle_per_molecule():
    mols = _mols()
    bp = BatchPlan([b for b, _ in mols])
    Js, Ks = bp.jk([D for _, D in mols], schwarz=1e-12)
    worst = 0.0
    for (b, D), J, K in zip(mols, Js, Ks):
        Jo, Ko = O.OracleBasis(b).jk(D, spherical=False, schwarz=1e-12)
        worst = max(worst, np.abs(J - Jo).max(), np.abs(K - Ko).max())
    st = bp.last_stats()
    print(f"{bp.n_mol} molecules, M {bp.M.min()}-{bp.M.max()}: max |d| vs oracle {worst:.2e}; {st}")
    assert worst < 1e-10
    assert st["shell_quartets"] > 0


def test_batch256_against_oracle_sketches():
    """The real 256-molecule batch (25-460 Cartesian functions) in one
    batched build at Schwarz 1e-12, against the oracle's J/K of every 16th
    molecule and the largest (tests/golden/batch256.npz): row sketches
    J X (X seeded N(0, 1), M x 4), diagonals and Frobenius norms.  An entry
    bound |dJ| <= 1e-10 implies |(dJ X)_ij| <= 1e-10 sum_k |X_kj|, the bound
    asserted here; a single entry off by 1e-7 already breaks it."""
    import os
    from test_oracle import basis_hash
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "batch256.npz"))
    mols = configs.batch256()
    bp = BatchPlan([b for b, _ in mols])
    Js, Ks = bp.jk([D for _, D in mols], schwarz=1e-12)
    worst = 0.0
    for i in g["pick"]:
        i = int(i)
        b, _ = mols[i]
        assert basis_hash(b) == str(g[f"sha{i}"])
        X = g[f"X{i}"]
        lim = 1e-10 * np.abs(X).sum(0)
        for W, w in ((Js[i], "J"), (Ks[i], "K")):
            d = np.abs(W @ X - g[f"{w}X{i}"])
            assert np.all(d <= lim), (i, w, (d / lim).max())
            assert np.abs(np.diag(W) - g[f"{w}d{i}"]).max() <= 1e-10, (i, w)
            assert abs(np.linalg.norm(W) - float(g[f"{w}f{i}"])) <= 1e-10 * W.shape[0], (i, w)
            worst = max(worst, (d / lim).max())
    print(f"batch256: {len(g['pick'])} molecules against the oracle, worst sketch error "
          f"{worst:.2e} of the 1e-10 entry bound")


def test_batch_matches_single_plans_and_shapes():
    mols = _mols(n=6, atoms=(8, 14), seed=5)
    bp = BatchPlan([b for b, _ in mols])
    Js, Ks = bp.jk([D for _, D in mols])
    singles = []
    for b, D in mols:
        singles.append(DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True).jk(D))
    for (J, K), Jb, Kb in zip(singles, Js, Ks):
        assert np.abs(J - Jb).max() < 1e-12 and np.abs(K - Kb).max() < 1e-12
    # every launch shape gives the same result
    for block, wpc in ((64, 4), (128, 16), (256, 32)):
        for X in range(6):
            for Y in range(X + 1):
                bp.set_shape(X, Y, block, wpc)
        J2, K2 = bp.jk([D for _, D in mols])
        for (J, K), Jb, Kb in zip(singles, J2, K2):
            assert np.abs(J - Jb).max() < 1e-12 and np.abs(K - Kb).max() < 1e-12
    rep = bp.autotune(repeats=2)
    print(rep)
    assert rep and all(v["block"] in (64, 128, 256) for v in rep.values())
    J3, K3 = bp.jk([D for _, D in mols])
    for (J, K), Jb, Kb in zip(singles, J3, K3):
        assert np.abs(J - Jb).max() < 1e-12 and np.abs(K - Kb).max() < 1e-12


def test_graph_replays_match_eager_builds():
    """Captured build replayed with new densities in the same buffer, eager
    work of another plan in between; every replay is checked."""
    mols = _mols(n=8, atoms=(3, 9), seed=11)
    bp = BatchPlan([b for b, _ in mols])
    other_b, other_D = configs.water7()
    other = DirectPlan(other_b, cartesian=True)
    dD = hipmem.DeviceArray((bp.total,))
    dJ = hipmem.DeviceArray((bp.total,))
    dK = hipmem.DeviceArray((bp.total,))
    rng = np.random.default_rng(3)
    for rep in range(4):
        dens = []
        for (b, D), M in zip(mols, bp.M):
            A = rng.standard_normal((M, M)) * 0.01
            dens.append(D + A + A.T)
        flat = bp.pack(dens)
        dD.upload(flat)
        bp.jk_device(dD, dJ, dK, use_graph=False)
        J_eager, K_eager = dJ.to_host(), dK.to_host()
        other.jk(other_D)  # unrelated eager work between replays
        dJ.upload(np.zeros(bp.total))
        bp.jk_device(dD, dJ, dK, use_graph=True)
        J_g, K_g = dJ.to_host(), dK.to_host()
        assert np.abs(J_g - J_eager).max() < 1e-12, rep
        assert np.abs(K_g - K_eager).max() < 1e-12, rep
