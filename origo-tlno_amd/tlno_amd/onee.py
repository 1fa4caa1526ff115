"""One-electron S/T/V over the HIP engine, mirroring
``tlno::integrals::onee`` (integrals/one_electron/onee_matrices.hpp:11-79):
``build_one_electron_matrices`` (Cartesian, component order) and
``build_one_electron_spherical_matrices`` (spherical M x M).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .basis import Basis
from .fock import DirectPlan, ScreeningOptions


@dataclass
class OneElectronMatrices:
    overlap: np.ndarray
    kinetic: np.ndarray
    nuclear: np.ndarray


def evaluate(plan: DirectPlan, charges, centers) -> OneElectronMatrices:
    Z = np.ascontiguousarray(charges, dtype=np.float64)
    R = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1)
    M = plan.M
    S = np.empty((M, M))
    T = np.empty((M, M))
    V = np.empty((M, M))
    _lib.check(_lib.load().tlno_gpu_onee(plan.handle, len(Z), Z.ctypes.data_as(_lib.D), R.ctypes.data_as(_lib.D),
                                         S.ctypes.data_as(_lib.D), T.ctypes.data_as(_lib.D),
                                         V.ctypes.data_as(_lib.D)))
    return OneElectronMatrices(S, T, V)


def build_one_electron_matrices(build: Basis) -> OneElectronMatrices:
    plan = DirectPlan(build, ScreeningOptions(), cartesian=True)
    return evaluate(plan, build.nuclear_charges, build.nuclear_centers)


def build_one_electron_spherical_matrices(build: Basis) -> OneElectronMatrices:
    plan = DirectPlan(build, ScreeningOptions(), cartesian=False)
    return evaluate(plan, build.nuclear_charges, build.nuclear_centers)
