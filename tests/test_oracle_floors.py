"""CPU: why the oracle-versus-reference comparisons of single integrals carry
a rounding floor (1e-13 Q_ab Q_cd for ERI blocks, 1e-13 sqrt|X_aa X_bb| for
S/T/V; tests/test_oracle.py).  On the organic5 entries that need it, the
40-digit values (tests/mp_integrals.py, McMurchie-Davidson in mpmath) show
that the oracle meets the strict bound |d| <= 1e-14 + 1e-12 |exact| and the
reference does not: the floor covers the reference's own rounding (entries
that vanish by symmetry come out as +-1e-13 there), not an error of this
build.  The GPU parity tests need no floor (tests/test_gpu_parity.py)."""
import os

import numpy as np

import mp_integrals as mpi
from oracle import oracle as O
from tlno_amd import configs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _strict(x, ex):
    return abs(x - ex) <= 1e-14 + 1e-12 * abs(ex)


def test_eri_floor_entries_are_reference_rounding():
    g = np.load(os.path.join(GOLD, "organic5.npz"))
    b, _ = configs.organic30(natoms=5)
    ob = O.OracleBasis(b)
    rows = []
    for k, q in enumerate(g["quads"]):
        q = [int(x) for x in q]
        ref = g["blocks"][g["block_off"][k]:g["block_off"][k + 1]]
        ours = ob.cart_block(*q).ravel()
        dims = [len(ob.groups[x][1]) for x in q]
        for e in np.nonzero(np.abs(ours - ref) > 1e-14 + 1e-12 * np.abs(ref))[0]:
            idx = np.unravel_index(e, dims)
            comps = [int(ob.groups[q[m]][1][idx[m]]) for m in range(4)]
            rows.append((abs(ours[e] - ref[e]), comps, ours[e], ref[e]))
    assert 0 < len(rows) < 0.1 * len(g["blocks"])
    rows.sort(key=lambda r: -r[0])
    pick = rows[:6] + [rows[i] for i in np.random.default_rng(5).choice(len(rows), 6, replace=False)]
    for _, comps, ours, ref in pick:
        ex = float(mpi.eri(b, *comps))
        assert _strict(ours, ex), (comps, ours, ex)
    # the largest differences are the reference's: it misses the strict bound
    assert sum(not _strict(r[3], float(mpi.eri(b, *r[1]))) for r in rows[:6]) == 6


def test_onee_floor_entries_are_reference_rounding():
    g = np.load(os.path.join(GOLD, "organic5.npz"))
    b, _ = configs.organic30(natoms=5)
    ob = O.OracleBasis(b)
    S, T, V = ob.onee(b.nuclear_charges, b.nuclear_centers, spherical=False)
    Z, R = b.nuclear_charges, b.nuclear_centers
    for name, x, r, f in (("S", S, g["S0"], lambda i, j: mpi.overlap(b, i, j)),
                          ("V", V, g["V0"], lambda i, j: mpi.nuclear(b, i, j, Z, R))):
        bad = np.argwhere(np.abs(x - r) > 1e-14 + 1e-12 * np.abs(r))
        assert len(bad) > 0
        order = sorted((tuple(int(v) for v in ij) for ij in bad), key=lambda ij: -abs(x[ij] - r[ij]))
        for ij in order[:6]:
            ex = float(f(*ij))
            assert _strict(x[ij], ex), (name, ij, x[ij], ex)
            assert not _strict(r[ij], ex), (name, ij, r[ij], ex)
