"""Learned shell-quartet screen: data, GPU training and plan integration.

The network (32 -> 64 -> 64 -> 1, ReLU, bf16 weights, FP32 sums) predicts
y = log10(max|(ab|cd)| / (Q_ab Q_cd)) <= 0 for a shell quartet from 32
features (pair exponents, reduced exponents, distances, contraction norms,
log Schwarz bounds, one-hot bra/ket classes).  Labels are computed by the
engine itself (tlno_gpu_quartet_maxabs), training runs on the GPU through the
library's own forward/backward/Adam kernels, and a build with the screen on
evaluates the network on the matrix cores for every candidate quartet.

Nothing here has a counterpart in the reference; the Schwarz-screened FP64
build is the screen's fallback and the judge of its J/K error.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib, configs
from .basis import Basis

NFEAT = 32
NPARAM = 6337
NALL = 6401
W_UNDER = 32.0  # loss weight of under-predicted magnitudes (the unsafe side)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "screen_mlp.npz")


def _fp(a):
    return a.ctypes.data_as(_lib.F)


class ScreenNetwork:
    """GPU-resident network with its Adam state (tlno_gpu_mlp_*)."""

    def __init__(self, seed: int = 0, params=None, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.tlno_gpu_mlp_create(int(device), C.c_uint(seed & 0xffffffff), C.byref(h)))
        self._h = h
        if params is not None:
            self.params = params

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().tlno_gpu_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(NALL, np.float32)
        _lib.check(_lib.load().tlno_gpu_mlp_get_params(self._h, _fp(out)))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (NALL,):
            raise ValueError(f"expected {NALL} parameters")
        _lib.check(_lib.load().tlno_gpu_mlp_set_params(self._h, _fp(p)))

    def set_normalization(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        _lib.check(_lib.load().tlno_gpu_mlp_set_normalization(self._h, len(X), _fp(X)))

    def loss_grad(self, X, t, w_under=W_UNDER):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_double()
        _lib.check(_lib.load().tlno_gpu_mlp_loss_grad(self._h, len(X), _fp(X), _fp(t), C.c_float(w_under), _fp(g),
                                                      C.byref(loss)))
        return loss.value, g

    def adam_step(self, X, t, lr, w_under=W_UNDER):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        _lib.check(_lib.load().tlno_gpu_mlp_adam_step(self._h, len(X), _fp(X), _fp(t), C.c_float(lr),
                                                      C.c_float(w_under)))

    def train(self, X, t, epochs=20, batch=1024, lr=2e-3, w_under=W_UNDER, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        hist = np.zeros(epochs, np.float32)
        _lib.check(_lib.load().tlno_gpu_mlp_train(self._h, len(X), _fp(X), _fp(t), int(epochs), int(batch),
                                                  C.c_float(lr), C.c_float(w_under), C.c_uint(seed & 0xffffffff),
                                                  _fp(hist)))
        return hist

    def predict(self, X, matrix_cores=True):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(len(X), np.float32)
        if len(X):
            _lib.check(_lib.load().tlno_gpu_mlp_predict(self._h, len(X), _fp(X), _fp(y), 1 if matrix_cores else 0))
        return y


def quartet_features(plan, P, Q) -> np.ndarray:
    """Features of quartets (P[i] | Q[i]) (plan pair indices), shape (n, 32)."""
    P = np.ascontiguousarray(P, dtype=np.int32)
    Q = np.ascontiguousarray(Q, dtype=np.int32)
    out = np.zeros((len(P), NFEAT), np.float32)
    _lib.check(_lib.load().tlno_gpu_quartet_features(plan.handle, len(P), P.ctypes.data_as(_lib.I),
                                                     Q.ctypes.data_as(_lib.I), _fp(out)))
    return out


def pair_classes(plan) -> np.ndarray:
    _, _, _, cb = plan.pairs()
    cls = np.zeros(plan.profile.pair_groups, np.int32)
    for c in range(len(cb) - 1):
        cls[cb[c]:cb[c + 1]] = c
    return cls


def sample_quartets(plan, n, rng, schwarz=1e-12):
    """n quartets drawn uniformly from the plan's Schwarz-surviving quartets,
    oriented as the builds orient them (bra class >= ket class, P >= Q within
    a class)."""
    _, _, Qb, _ = plan.pairs()
    cls = pair_classes(plan)
    npair = len(Qb)
    Ps, Qs = [], []
    need = n
    while need > 0:
        m = max(4 * need, 1024)
        a = rng.integers(0, npair, m)
        b = rng.integers(0, npair, m)
        swap = (cls[a] < cls[b]) | ((cls[a] == cls[b]) & (a < b))
        P = np.where(swap, b, a)
        Q = np.where(swap, a, b)
        ok = Qb[P] * Qb[Q] >= schwarz
        Ps.append(P[ok][:need])
        Qs.append(Q[ok][:need])
        need -= len(Ps[-1])
    return np.concatenate(Ps).astype(np.int32), np.concatenate(Qs).astype(np.int32)


def labels(plan, P, Q) -> np.ndarray:
    """log10(max|(ab|cd)| / (Q_ab Q_cd)) computed by the engine (float64)."""
    _, _, Qb, _ = plan.pairs()
    m = plan.quartet_maxabs(P, Q)
    return np.log10(np.maximum(m, 1e-300)) - np.log10(Qb[P] * Qb[Q])


MIN_SEPARATION_A = 0.85  # covers bonded distances such as O-H 0.96 A
ZERO_LABEL = -10.0  # quartets this far below their Schwarz bound are symmetry zeros


def training_molecule(rng, natoms):
    """A molecule from the batch generator (heavy 3s2p1d / H 2s1p, three
    primitives per shell, exponents log-uniform in [0.1, 150]) with a basis
    set drawn for this molecule alone, so that the training set spans many
    basis sets, and atoms at least MIN_SEPARATION_A apart."""
    eb = configs._element_bases(rng, ["C", "N", "O", "H"], [0, 0, 0, 1, 1, 2], [0, 0, 1], (3, 3), (0.1, 150.0))
    box = 12.0 * (natoms / 30.0) ** (1.0 / 3.0)
    atoms = configs._random_atoms(rng, natoms, box, MIN_SEPARATION_A)
    return configs._build(atoms, eb, f"train{natoms}")


@dataclass
class TrainingSet:
    X: np.ndarray  # (n, 32) float32
    t: np.ndarray  # (n,) float64 labels
    molecule: np.ndarray  # (n,) source molecule index

    def trainable(self):
        """Mask of the samples the network is fitted to: symmetry zeros
        (label below ZERO_LABEL, e.g. odd-parity one-centre quartets) are left
        out; their -300 labels would put a cliff into a smooth function, and
        a screen that keeps them only loses a negligible amount of work."""
        return self.t >= ZERO_LABEL


def make_training_set(n_molecules=500, per_molecule=4000, seed=configs.SEED0 + 6, atoms=(3, 40), progress=None):
    """Features and engine-computed labels of per_molecule quartets from each
    of n_molecules generated molecules (3-40 atoms, uniform)."""
    from .fock import DirectPlan, ScreeningOptions
    rng = np.random.default_rng(seed)
    Xs, ts, ms = [], [], []
    for i in range(n_molecules):
        na = int(rng.integers(atoms[0], atoms[1] + 1))
        b = training_molecule(rng, na)
        plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
        P, Q = sample_quartets(plan, per_molecule, rng)
        Xs.append(quartet_features(plan, P, Q))
        ts.append(labels(plan, P, Q))
        ms.append(np.full(len(P), i, np.int32))
        plan.close()
        if progress and (i + 1) % 50 == 0:
            progress(i + 1)
    return TrainingSet(np.concatenate(Xs), np.concatenate(ts), np.concatenate(ms))


def save_weights(path, params, margin, **meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    arrays = {"params": np.asarray(params, np.float32), "margin": np.float64(margin)}
    for k, v in meta.items():
        arrays[k] = np.asarray(v)
    np.savez(path, **arrays)


def load_weights(path=DEFAULT_WEIGHTS):
    """(params float32[6401], margin) of a trained network."""
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), float(z["margin"])


def choose_margin(residual, quantile=0.9999, pad=0.25):
    """Safety offset (decades) from held-out residuals t - y: the given
    quantile of the under-prediction plus a fixed pad."""
    return float(np.quantile(residual, quantile) + pad)
