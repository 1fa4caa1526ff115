"""Seeded synthetic workloads (molecules, contracted Cartesian bases, densities).

Five workloads, all coordinates converted from angstrom to bohr
(1 bohr = 0.529177210903 A, the unit of the reference's test molecules,
test_molecules/README.md:3):

  water7     water, O at origin, H at (+-0.7570, 0.5859, 0) A; O 2s1p, H 1s,
             3 primitives per shell, exponents log-uniform in [0.1, 150],
             raw coefficients U(0.1, 1); D = C C^T, C 7x5 N(0,1)/sqrt(7).
  organic30  30 atoms uniform in a 12 A cube, minimum separation 1.0 A,
             C/N/O/H with probabilities .3/.1/.1/.5; heavy 3s2p1d, H 2s1p
             (Cartesian d); 1-6 primitives per shell, exponents log-uniform in
             [0.08, 5000]; D = C C^T, C nbf x 60 N(0,1), ||D||_F = 1;
             Schwarz threshold 1e-12.
  batch256   256 molecules of 3-40 atoms (uniform), placed like organic30 at
             the same number density, water7-style 3-primitive shells with
             heavy 3s2p1d / H 2s1p.
  water40    40 rigid waters (O-H 0.96 A, 104.5 deg), random orientations in an
             11 A cube, O-O >= 2.5 A; O 3s2p1d, H 2s1p (1000 functions);
             exponents as water7.
  water100   100 waters laid out like water40 at the same number density
             (about 2500 functions).
  organic_f  organic30's generator with an f polarization shell (1-2
             primitives, exponents in [0.3, 6]) on the heavy atoms, a few
             atoms: the f-shell parity workload.
  organic_g  the same with an f and a g polarization shell: the g-shell
             parity workload.

Bases are drawn once per element (a synthetic basis SET) from the workload's
seed, so equal elements share contractions as a real basis set does.  Seeds
are 20261203 + workload number.
"""
from __future__ import annotations

import numpy as np

from .basis import BOHR_PER_ANGSTROM, Basis

SEED0 = 20261203
Z_OF = {"H": 1, "C": 6, "N": 7, "O": 8}


def _draw_shells(rng, spec, nprim_range, exp_range):
    """spec: list of L values; returns [(L, exps, raw_coefs)]"""
    out = []
    lo, hi = np.log(exp_range[0]), np.log(exp_range[1])
    for L in spec:
        npr = int(rng.integers(nprim_range[0], nprim_range[1] + 1))
        e = np.sort(np.exp(rng.uniform(lo, hi, npr)))[::-1].copy()
        c = rng.uniform(0.1, 1.0, npr)
        out.append((L, e, c))
    return out


def _element_bases(rng, elements, heavy_spec, h_spec, nprim_range, exp_range):
    bases = {}
    for el in elements:
        spec = h_spec if el == "H" else heavy_spec
        bases[el] = _draw_shells(rng, spec, nprim_range, exp_range)
    return bases


def _build(atoms, ebases, name):
    """atoms: list of (element, xyz_bohr)"""
    shells, Z, R = [], [], []
    for el, xyz in atoms:
        Z.append(Z_OF[el])
        R.append(xyz)
        for L, e, c in ebases[el]:
            shells.append((np.asarray(xyz, dtype=np.float64), L, e, c))
    return Basis.from_shells(shells, Z, R, name=name)


def _density(rng, nbf, nocc, scale=None):
    Cm = rng.standard_normal((nbf, nocc))
    if scale is not None:
        Cm = Cm * scale
    D = Cm @ Cm.T
    return D


def _random_atoms(rng, natoms, box_A, min_sep_A, probs=(0.3, 0.1, 0.1, 0.5)):
    els = np.array(["C", "N", "O", "H"])
    pos = []
    tries = 0
    while len(pos) < natoms:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place atoms")
        p = rng.uniform(0.0, box_A, 3)
        if all(np.linalg.norm(p - q) >= min_sep_A for q in pos):
            pos.append(p)
    el = rng.choice(els, size=natoms, p=list(probs))
    return [(str(e), p * BOHR_PER_ANGSTROM) for e, p in zip(el, pos)]


def spherical_density(M, seed, nocc):
    """D = C C^T / ||C C^T||_F with C an M x nocc N(0,1) matrix (for the
    spherical-basis builds of a workload)."""
    rng = np.random.default_rng(seed)
    Cm = rng.standard_normal((M, nocc))
    D = Cm @ Cm.T
    return D / np.linalg.norm(D)


def water7(seed=SEED0 + 1):
    rng = np.random.default_rng(seed)
    eb = {"O": _draw_shells(rng, [0, 0, 1], (3, 3), (0.1, 150.0)),
          "H": _draw_shells(rng, [0], (3, 3), (0.1, 150.0))}
    atoms = [("O", np.zeros(3)),
             ("H", np.array([0.7570, 0.5859, 0.0]) * BOHR_PER_ANGSTROM),
             ("H", np.array([-0.7570, 0.5859, 0.0]) * BOHR_PER_ANGSTROM)]
    b = _build(atoms, eb, "water7")
    Cm = rng.standard_normal((b.n, 5)) / np.sqrt(7.0)
    return b, Cm @ Cm.T


def organic30(seed=SEED0 + 2, natoms=30, box_A=12.0):
    rng = np.random.default_rng(seed)
    atoms = _random_atoms(rng, natoms, box_A, 1.0)
    eb = _element_bases(rng, ["C", "N", "O", "H"], [0, 0, 0, 1, 1, 2], [0, 0, 1],
                        (1, 6), (0.08, 5000.0))
    b = _build(atoms, eb, f"organic{natoms}")
    D = _density(rng, b.n, 60)
    D /= np.linalg.norm(D)
    return b, D


def batch256(seed=SEED0 + 3, nmol=256, atom_range=(3, 40)):
    rng = np.random.default_rng(seed)
    eb = _element_bases(rng, ["C", "N", "O", "H"], [0, 0, 0, 1, 1, 2], [0, 0, 1],
                        (3, 3), (0.1, 150.0))
    mols = []
    for i in range(nmol):
        na = int(rng.integers(atom_range[0], atom_range[1] + 1))
        box = 12.0 * (na / 30.0) ** (1.0 / 3.0)
        atoms = _random_atoms(rng, na, box, 1.0)
        b = _build(atoms, eb, f"batch{i}")
        D = _density(rng, b.n, max(1, b.n // 5))
        D /= np.linalg.norm(D)
        mols.append((b, D))
    return mols


def _waters(rng, nwat, box_A, ebases, name):
    r_oh, ang = 0.96, np.deg2rad(104.5)
    local = np.array([[0.0, 0.0, 0.0],
                      [r_oh * np.sin(ang / 2), r_oh * np.cos(ang / 2), 0.0],
                      [-r_oh * np.sin(ang / 2), r_oh * np.cos(ang / 2), 0.0]])
    centers = []
    tries = 0
    while len(centers) < nwat:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place waters")
        p = rng.uniform(0.0, box_A, 3)
        if all(np.linalg.norm(p - q) >= 2.5 for q in centers):
            centers.append(p)
    atoms = []
    for c in centers:
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        xyz = local @ q.T + c
        for el, x in zip(("O", "H", "H"), xyz):
            atoms.append((el, x * BOHR_PER_ANGSTROM))
    return _build(atoms, ebases, name)


def water40(seed=SEED0 + 4, nwat=40, box_A=11.0):
    rng = np.random.default_rng(seed)
    eb = _element_bases(rng, ["O", "H"], [0, 0, 0, 1, 1, 2], [0, 0, 1], (3, 3), (0.1, 150.0))
    b = _waters(rng, nwat, box_A, eb, f"water{nwat}")
    D = _density(rng, b.n, 5 * nwat)
    D /= np.linalg.norm(D)
    return b, D


def water100(seed=SEED0 + 5, nwat=100):
    box = 11.0 * (nwat / 40.0) ** (1.0 / 3.0)
    return water40(seed=seed, nwat=nwat, box_A=box)


def water_cluster(nwat, seed=SEED0 + 4):
    """water40-style cluster of any size at water40's number density."""
    box = 11.0 * (nwat / 40.0) ** (1.0 / 3.0)
    return water40(seed=seed, nwat=nwat, box_A=box)


def organic_f(seed=SEED0 + 8, natoms=5, box_A=4.5):
    """organic30's generator with f shells (the reference's L <= 4 range
    beyond d): heavy atoms 3s2p1d as organic30 (1-6 primitives, exponents
    log-uniform in [0.08, 5000]) plus one f polarization shell of 1-2
    primitives with exponents log-uniform in [0.3, 6] (the range of the f
    shells of real basis sets, e.g. def2-TZVP's), H 2s1p; elements C/N/O/H
    with probabilities .4/.2/.2/.2, minimum separation 1.0 A; D = C C^T with
    C nbf x nbf/5 N(0,1), ||D||_F = 1."""
    rng = np.random.default_rng(seed)
    atoms = _random_atoms(rng, natoms, box_A, 1.0, probs=(0.4, 0.2, 0.2, 0.2))
    eb = _element_bases(rng, ["C", "N", "O", "H"], [0, 0, 0, 1, 1, 2], [0, 0, 1],
                        (1, 6), (0.08, 5000.0))
    for el in ("C", "N", "O"):
        eb[el] = eb[el] + _draw_shells(rng, [3], (1, 2), (0.3, 6.0))
    b = _build(atoms, eb, f"organic_f{natoms}")
    D = _density(rng, b.n, max(1, b.n // 5))
    D /= np.linalg.norm(D)
    return b, D


def organic_g(seed=SEED0 + 9, natoms=3, box_A=3.5):
    """organic_f with a g polarization shell as well: heavy atoms 3s2p1d1f1g
    (the f and g shells 1-2 primitives, exponents log-uniform in [0.3, 6]),
    H 2s1p, a few atoms: the g-shell parity workload (the reference's
    largest angular momentum, L = 4)."""
    rng = np.random.default_rng(seed)
    atoms = _random_atoms(rng, natoms, box_A, 1.0, probs=(0.4, 0.2, 0.2, 0.2))
    eb = _element_bases(rng, ["C", "N", "O", "H"], [0, 0, 0, 1, 1, 2], [0, 0, 1],
                        (1, 6), (0.08, 5000.0))
    for el in ("C", "N", "O"):
        eb[el] = eb[el] + _draw_shells(rng, [3, 4], (1, 2), (0.3, 6.0))
    b = _build(atoms, eb, f"organic_g{natoms}")
    D = _density(rng, b.n, max(1, b.n // 5))
    D /= np.linalg.norm(D)
    return b, D


def distinct_pairs():
    """Three centres (He, H, H charges) whose five shells (s with 1
    primitive, p with 2, s with 3, s with 4, d with 1) give every shell pair
    its own (L_a, L_b, primitive-pair count) class: no two pairs share a
    bra packet of a SIMD engine that groups pairs by that key, so a
    per-primitive-quartet screen and a per-packet one select the same
    primitive quartets.  Spherical density: symmetric N(0, 1) entries."""
    from .basis import Basis
    A = np.array([0.0, 0.0, 0.0])
    B = np.array([0.0, 0.9, 1.3])
    C = np.array([1.4, -0.3, 0.2])
    shells = [(A, 0, [3.1], [1.0]), (A, 1, [5.0, 0.7], [0.4, 0.8]), (B, 0, [12.0, 2.1, 0.4], [0.2, 0.5, 0.6]),
              (C, 0, [8.0, 1.1, 0.3, 0.09], [0.1, 0.3, 0.5, 0.4]), (C, 2, [1.9], [1.0])]
    b = Basis.from_shells(shells, [2, 1, 1], [A, B, C], name="distinct_pairs")
    X = np.random.default_rng(SEED0 + 7).standard_normal((b.nsph(), b.nsph()))
    return b, X + X.T
