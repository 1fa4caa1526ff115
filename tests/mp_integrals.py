"""40-digit reference values of single integrals (test infrastructure): the
McMurchie-Davidson Hermite expansion evaluated in mpmath.

  E^{ij}_t   1D Hermite coefficients, E^{00}_0 = exp(-mu X_AB^2),
             E^{i+1,j}_t = E^{ij}_{t-1}/(2p) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
             (the same with X_PB for j);
  R^n_{tuv}  R^n_{000} = (-2 alpha)^n F_n(alpha R^2),
             R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}, likewise u, v;
  (ab|cd)    2 pi^{5/2} / (p q sqrt(p+q)) sum_{tuv} E^{ab}_{tuv}
             sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau, u+nu, v+phi}
  S          (pi/p)^{3/2} E_0 E_0 E_0;  V = -Z 2 pi/p sum_{tuv} E_{tuv} R_{tuv}(p, P - C).
F_n(T) = gamma(n + 1/2, 0..T) / (2 T^{n+1/2}) from mpmath's incomplete gamma.
These are textbook expressions independent of the Rys quadrature and the
Obara-Saika recurrences both the engine and the oracle use."""
import mpmath as mp
import numpy as np

mp.mp.dps = 40


def boys(n, T):
    if T == 0:
        return mp.mpf(1) / (2 * n + 1)
    return mp.gammainc(mp.mpf(n) + mp.mpf(1) / 2, 0, T) / (2 * T ** (mp.mpf(n) + mp.mpf(1) / 2))


def hermite_E(i, j, a, b, A, B):
    """E[t] for t = 0..i+j along one axis."""
    p = a + b
    mu = a * b / p
    X = A - B
    P = (a * A + b * B) / p
    xpa, xpb = P - A, P - B
    cur = {(0, 0): [mp.e ** (-mu * X * X)]}

    def step(prev, x):
        n = len(prev)
        out = []
        for t in range(n + 1):
            v = mp.mpf(0)
            if t >= 1:
                v += prev[t - 1] / (2 * p)
            if t < n:
                v += x * prev[t]
            if t + 1 < n:
                v += (t + 1) * prev[t + 1]
            out.append(v)
        return out

    row = cur[(0, 0)]
    for _ in range(i):
        row = step(row, xpa)
    for _ in range(j):
        row = step(row, xpb)
    return row


def hermite_R(tmax, umax, vmax, alpha, PQ):
    T = alpha * (PQ[0] ** 2 + PQ[1] ** 2 + PQ[2] ** 2)
    nmax = tmax + umax + vmax
    F = [boys(n, T) * (-2 * alpha) ** n for n in range(nmax + 1)]
    memo = {}

    def R(n, t, u, v):
        if t < 0 or u < 0 or v < 0:
            return mp.mpf(0)
        key = (n, t, u, v)
        if key in memo:
            return memo[key]
        if t > 0:
            r = (t - 1) * R(n + 1, t - 2, u, v) + PQ[0] * R(n + 1, t - 1, u, v)
        elif u > 0:
            r = (u - 1) * R(n + 1, t, u - 2, v) + PQ[1] * R(n + 1, t, u - 1, v)
        elif v > 0:
            r = (v - 1) * R(n + 1, t, u, v - 2) + PQ[2] * R(n + 1, t, u, v - 1)
        else:
            r = F[n]
        memo[key] = r
        return r

    return lambda t, u, v: R(0, t, u, v)


def _prims(b, comp):
    off = int(np.sum(b.nprim[:comp]))
    n = int(b.nprim[comp])
    return ([mp.mpf(float(x)) for x in b.exps[off:off + n]], [mp.mpf(float(x)) for x in b.coefs[off:off + n]],
            [mp.mpf(float(x)) for x in b.centers[comp]], [int(x) for x in b.ang[comp]])


def eri(b, ca, cb, cc, cd):
    """(ca cb | cc cd) over the stored (normalized) coefficients, components
    given by basis index."""
    ea, wa, A, la = _prims(b, ca)
    eb, wb, B, lb = _prims(b, cb)
    ec, wc, C, lc = _prims(b, cc)
    ed, wd, D, ld = _prims(b, cd)
    tot = mp.mpf(0)
    for a, ka in zip(ea, wa):
        for bb, kb in zip(eb, wb):
            p = a + bb
            P = [(a * A[x] + bb * B[x]) / p for x in range(3)]
            Eab = [hermite_E(la[x], lb[x], a, bb, A[x], B[x]) for x in range(3)]
            for c, kc in zip(ec, wc):
                for d, kd in zip(ed, wd):
                    q = c + d
                    Qc = [(c * C[x] + d * D[x]) / q for x in range(3)]
                    Ecd = [hermite_E(lc[x], ld[x], c, d, C[x], D[x]) for x in range(3)]
                    alpha = p * q / (p + q)
                    R = hermite_R(len(Eab[0]) + len(Ecd[0]), len(Eab[1]) + len(Ecd[1]),
                                  len(Eab[2]) + len(Ecd[2]), alpha, [P[x] - Qc[x] for x in range(3)])
                    s = mp.mpf(0)
                    for t, ex in enumerate(Eab[0]):
                        for u, ey in enumerate(Eab[1]):
                            for v, ez in enumerate(Eab[2]):
                                e1 = ex * ey * ez
                                if e1 == 0:
                                    continue
                                for tau, fx in enumerate(Ecd[0]):
                                    for nu, fy in enumerate(Ecd[1]):
                                        for phi, fz in enumerate(Ecd[2]):
                                            sgn = -1 if (tau + nu + phi) % 2 else 1
                                            s += e1 * sgn * fx * fy * fz * R(t + tau, u + nu, v + phi)
                    tot += ka * kb * kc * kd * 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) * s
    return tot


def overlap(b, ca, cb):
    ea, wa, A, la = _prims(b, ca)
    eb, wb, B, lb = _prims(b, cb)
    tot = mp.mpf(0)
    for a, ka in zip(ea, wa):
        for bb, kb in zip(eb, wb):
            p = a + bb
            e = [hermite_E(la[x], lb[x], a, bb, A[x], B[x])[0] for x in range(3)]
            tot += ka * kb * (mp.pi / p) ** mp.mpf(1.5) * e[0] * e[1] * e[2]
    return tot


def nuclear(b, ca, cb, Z, Rn):
    ea, wa, A, la = _prims(b, ca)
    eb, wb, B, lb = _prims(b, cb)
    tot = mp.mpf(0)
    for a, ka in zip(ea, wa):
        for bb, kb in zip(eb, wb):
            p = a + bb
            P = [(a * A[x] + bb * B[x]) / p for x in range(3)]
            E = [hermite_E(la[x], lb[x], a, bb, A[x], B[x]) for x in range(3)]
            for z, c in zip(Z, Rn):
                C = [mp.mpf(float(x)) for x in c]
                R = hermite_R(len(E[0]), len(E[1]), len(E[2]), p, [P[x] - C[x] for x in range(3)])
                s = mp.mpf(0)
                for t, ex in enumerate(E[0]):
                    for u, ey in enumerate(E[1]):
                        for v, ez in enumerate(E[2]):
                            s += ex * ey * ez * R(t, u, v)
                tot += -mp.mpf(int(z)) * 2 * mp.pi / p * ka * kb * s
    return tot


def kinetic(b, ca, cb):
    """-1/2 <a|nabla^2|b> in the symmetric form 1/2 <grad a|grad b>, per axis
    T_x = 1/2 [4ab S(i+1,j+1) - 2ib S(i-1,j+1) - 2ja S(i+1,j-1) + ij S(i-1,j-1)]
    with S(i, j) the 1D overlap sqrt(pi/p) E^{ij}_0."""
    ea, wa, A, la = _prims(b, ca)
    eb, wb, B, lb = _prims(b, cb)
    tot = mp.mpf(0)
    for a, ka in zip(ea, wa):
        for bb, kb in zip(eb, wb):
            p = a + bb

            def s1(i, j, x):
                if i < 0 or j < 0:
                    return mp.mpf(0)
                return mp.sqrt(mp.pi / p) * hermite_E(i, j, a, bb, A[x], B[x])[0]
            s, t = [], []
            for x in range(3):
                i, j = la[x], lb[x]
                tv = 4 * a * bb * s1(i + 1, j + 1, x) - 2 * i * bb * s1(i - 1, j + 1, x) \
                    - 2 * j * a * s1(i + 1, j - 1, x) + i * j * s1(i - 1, j - 1, x)
                s.append(s1(i, j, x))
                t.append(tv / 2)
            tot += ka * kb * (t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2])
    return tot
