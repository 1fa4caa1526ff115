"""Algorithmic flop model of the Rys ERI + J/K path (one function, used by the
benchmark and the roofline report).

Per primitive quartet of class (la lb|lc ld), N = (la+lb+lc+ld)//2 + 1 roots,
LAB = la+lb, LCD = lc+ld, nc = ncart(la) ncart(lb) ncart(lc) ncart(ld):

  F_prim = 30                                    pair-pair setup (T, prefactor)
         + 38 N                                  roots, weights, B00/B10/B01, C00/D00
         + 3 N [5((LAB+1)(LCD+1) - 1) + 2 HB + 2 HK]   2D VRR + bra/ket HRR per axis
         + 3 N nc                                triple product, 2 mul + 1 add

  HB = sum_{j=1..lb} (LAB - j + 1)(LCD + 1)
  HK = sum_{l=1..ld} (la + 1)(lb + 1)(LCD - l + 1)

Per contracted shell quartet the J+K digestion adds 12 flops per output
integral (J only 4, K only 8).  Worked values: (ss|ss) 71, (ps|ss) 92,
(ps|ps) 250, (ds|ds) 828, (pp|pp) 1485, (dd|dd) 24400.
"""
from __future__ import annotations

CLASS_L = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (3, 0), (3, 1), (3, 2), (3, 3),
           (4, 0), (4, 1), (4, 2), (4, 3), (4, 4)]
CLASS_NAME = ["ss", "ps", "pp", "ds", "dp", "dd", "fs", "fp", "fd", "ff", "gs", "gp", "gd", "gf", "gg"]


def ncart(L):
    return (L + 1) * (L + 2) // 2


def f_prim(la, lb, lc, ld):
    LAB, LCD = la + lb, lc + ld
    N = (LAB + LCD) // 2 + 1
    nc = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    HB = sum((LAB - j + 1) * (LCD + 1) for j in range(1, lb + 1))
    HK = sum((la + 1) * (lb + 1) * (LCD - l + 1) for l in range(1, ld + 1))
    return 30 + 38 * N + 3 * N * (5 * ((LAB + 1) * (LCD + 1) - 1) + 2 * HB + 2 * HK) + 3 * N * nc


def n_out(la, lb, lc, ld):
    return ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)


def class_flops(X, Y, prim_quartets, shell_quartets, digest_per_integral=12):
    la, lb = CLASS_L[X]
    lc, ld = CLASS_L[Y]
    return f_prim(la, lb, lc, ld) * prim_quartets + digest_per_integral * n_out(la, lb, lc, ld) * shell_quartets
