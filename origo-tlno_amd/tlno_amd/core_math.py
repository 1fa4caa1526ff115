"""Boys function and Rys rules evaluated by the device evaluators used in the
kernels; mirrors ``tlno::core::math`` (core/math/general_math.hpp:6-22) and
the Rys rule of integrals/two_electron/rys/rys_engine.ipp:3820-3837."""
from __future__ import annotations

import numpy as np

from . import _lib


def boys(T, mmax):
    """F_0..F_mmax(T) for an array of T: shape (n, mmax+1)."""
    T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
    out = np.empty((len(T), mmax + 1))
    _lib.check(_lib.load().tlno_gpu_boys(len(T), T.ctypes.data_as(_lib.D), int(mmax), out.ctypes.data_as(_lib.D)))
    return out


def boys_single(order, x):
    return float(boys([x], order)[0, order])


def boys_r000_all(max_n, x, p):
    """(-2p)^n F_n(x), n = 0..max_n (general_math.hpp:7)."""
    F = boys([x], max_n)[0]
    return F * (-2.0 * p) ** np.arange(max_n + 1)


def rys(N, T):
    """roots t^2 and weights, each shape (n, N)."""
    T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
    r = np.empty((len(T), N))
    w = np.empty((len(T), N))
    _lib.check(_lib.load().tlno_gpu_rys(int(N), len(T), T.ctypes.data_as(_lib.D), r.ctypes.data_as(_lib.D),
                                        w.ctypes.data_as(_lib.D)))
    return r, w
