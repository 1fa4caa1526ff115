"""Basis data model, mirroring the reference's types.

The reference's unit of basis data is one Cartesian COMPONENT
(``tlno::core::basis::Shell``, core/basis/types.hpp:31-36): a centre, the
powers (lx, ly, lz) and its primitives with stored, normalized coefficients.
``Basis`` keeps exactly that layout as flat numpy arrays (the drop-in C ABI
takes the same arrays), plus the nuclear charges/centres of
``BasisBuildResult`` (types.hpp:38-43).

``cartesian_tuples`` follows core/basis/cartesian_shells.hpp:16-28
(descending x, then y).  ``normalize_contracted`` restates
core/shell/shell_contract.cpp:9-48 with core/shell/shell_math.cpp:50-77:
primitive normalization times a contraction scale giving every component
unit self-overlap.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903


def cartesian_tuples(L):
    out = []
    for x in range(L, -1, -1):
        for y in range(L - x, -1, -1):
            out.append((x, y, L - x - y))
    return out


def ncart(L):
    return (L + 1) * (L + 2) // 2


def _dfact(n):
    """(2n-1)!!"""
    v = 1.0
    k = 2 * n - 1
    while k > 0:
        v *= k
        k -= 2
    return v


def normalize_contracted(exps, coefs, ang):
    exps = np.asarray(exps, dtype=np.float64)
    coefs = np.asarray(coefs, dtype=np.float64)
    L = int(sum(ang))
    df = _dfact(ang[0]) * _dfact(ang[1]) * _dfact(ang[2])
    c = coefs * np.sqrt(2.0 ** (2 * L + 1.5) * exps ** (L + 1.5) / (np.pi ** 1.5 * df))
    p = exps[:, None] + exps[None, :]
    ov = (np.pi / p) ** 1.5
    for a in ang:
        ov = ov * _dfact(a) / (2.0 * p) ** a
    n2 = float(c @ ov @ c)
    return c / np.sqrt(n2)


@dataclass
class Basis:
    centers: np.ndarray          # (n, 3) bohr
    ang: np.ndarray              # (n, 3) int32
    nprim: np.ndarray            # (n,) int32
    exps: np.ndarray             # (sum nprim,)
    coefs: np.ndarray            # (sum nprim,) stored (normalized) coefficients
    nuclear_charges: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    nuclear_centers: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    name: str = ""

    @property
    def n(self):
        return len(self.nprim)

    @classmethod
    def from_shells(cls, shells, nuclear_charges=(), nuclear_centers=(), name=""):
        """shells: iterable of (center_bohr[3], L, exps, raw_coefs)."""
        centers, ang, nprim, exps, coefs = [], [], [], [], []
        for center, L, e, c in shells:
            e = np.asarray(e, dtype=np.float64)
            c = np.asarray(c, dtype=np.float64)
            for t in cartesian_tuples(L):
                centers.append(center)
                ang.append(t)
                nprim.append(len(e))
                exps.append(e)
                coefs.append(normalize_contracted(e, c, t))
        return cls(np.asarray(centers, dtype=np.float64).reshape(-1, 3),
                   np.asarray(ang, dtype=np.int32).reshape(-1, 3),
                   np.asarray(nprim, dtype=np.int32),
                   np.concatenate(exps) if exps else np.zeros(0),
                   np.concatenate(coefs) if coefs else np.zeros(0),
                   np.asarray(nuclear_charges, dtype=np.int32),
                   np.asarray(nuclear_centers, dtype=np.float64).reshape(-1, 3),
                   name)

    def shell_L(self):
        return self.ang.sum(axis=1)

    def nsph(self):
        """spherical function count: 2L+1 per complete Cartesian shell."""
        L = self.shell_L()
        return int(round(sum((2 * l + 1) / ncart(l) for l in L)))
