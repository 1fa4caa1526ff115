"""Molecule batches: many molecules' J/K builds in one set of persistent
class launches, optionally replayed from a captured HIP graph, with launch
shapes from the autotuner (C ABI tlno_gpu_batch_*; no reference
counterpart: the reference builds one molecule per call).

D, J and K of molecule m are M_m x M_m Cartesian blocks (component order of
the molecule's basis) packed at offsets[m] of one flat array."""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _lib
from .basis import Basis

CLASS_NAMES = ["ss", "ps", "pp", "ds", "dp", "dd"]


class BatchPlan:
    def __init__(self, bases: list[Basis], device: int = 0):
        if not bases:
            raise ValueError("empty batch")
        lib = _lib.load()
        _lib.check(lib.tlno_gpu_init(int(device)))
        cat = lambda key, dt: np.ascontiguousarray(np.concatenate([np.asarray(getattr(b, key)).reshape(-1)
                                                                   for b in bases]), dtype=dt)
        self._arrays = (np.array([len(b.nprim) for b in bases], np.int32), cat("centers", np.float64),
                        cat("ang", np.int32), cat("nprim", np.int32), cat("exps", np.float64),
                        cat("coefs", np.float64))
        nc, c, a, n, e, f = self._arrays
        h = C.c_void_p()
        _lib.check(lib.tlno_gpu_batch_create(len(bases), nc.ctypes.data_as(_lib.I), c.ctypes.data_as(_lib.D),
                                             a.ctypes.data_as(_lib.I), n.ctypes.data_as(_lib.I),
                                             e.ctypes.data_as(_lib.D), f.ctypes.data_as(_lib.D), C.byref(h)))
        self._h = h
        nm = C.c_int()
        tot = C.c_longlong()
        self.M = np.zeros(len(bases), np.int32)
        self.offsets = np.zeros(len(bases), np.int64)
        _lib.check(lib.tlno_gpu_batch_info(h, C.byref(nm), C.byref(tot), self.M.ctypes.data_as(_lib.I),
                                           self.offsets.ctypes.data_as(_lib.LL)))
        self.n_mol = nm.value
        self.total = tot.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().tlno_gpu_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, mats) -> np.ndarray:
        out = np.empty(self.total)
        for m, X in enumerate(mats):
            X = np.asarray(X, dtype=np.float64)
            if X.shape != (self.M[m], self.M[m]):
                raise ValueError(f"molecule {m}: expected {self.M[m]} x {self.M[m]}")
            out[self.offsets[m]:self.offsets[m] + X.size] = X.ravel()
        return out

    def unpack(self, flat):
        return [np.asarray(flat[o:o + M * M]).reshape(M, M) for o, M in zip(self.offsets, self.M)]

    def jk(self, densities, schwarz=1e-12):
        """Per-molecule (J, K) lists from host densities."""
        D = self.pack(densities)
        J = np.empty(self.total)
        K = np.empty(self.total)
        _lib.check(_lib.load().tlno_gpu_batch_build_jk(self._h, D.ctypes.data_as(_lib.D), J.ctypes.data_as(_lib.D),
                                                       K.ctypes.data_as(_lib.D), float(schwarz)))
        return self.unpack(J), self.unpack(K)

    def jk_device(self, dD, dJ, dK, schwarz=1e-12, use_graph=False):
        """Build from device buffers (hipmem.DeviceArray of `total` doubles)."""
        _lib.check(_lib.load().tlno_gpu_batch_build_jk_device(self._h, dD.ptr, dJ.ptr if dJ else None,
                                                              dK.ptr if dK else None, float(schwarz),
                                                              1 if use_graph else 0))

    def autotune(self, schwarz=1e-12, repeats=3) -> dict:
        buf = C.create_string_buffer(8192)
        _lib.check(_lib.load().tlno_gpu_batch_autotune(self._h, float(schwarz), int(repeats), buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_shape(self, bra_class, ket_class, block, waves_per_cu):
        _lib.check(_lib.load().tlno_gpu_batch_set_shape(self._h, int(bra_class), int(ket_class), int(block),
                                                        int(waves_per_cu)))

    def last_stats(self):
        sq = C.c_longlong()
        pq = C.c_longlong()
        ms = C.c_double()
        _lib.check(_lib.load().tlno_gpu_batch_last_stats(self._h, C.byref(sq), C.byref(pq), C.byref(ms)))
        return {"shell_quartets": sq.value, "prim_quartets": pq.value, "ms": ms.value}
