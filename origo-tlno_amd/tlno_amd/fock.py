"""Direct J/K/G builders over the HIP engine, mirroring the reference's
``tlno::integrals::twoe::fock`` and ``tlno::integrals::twoe`` interfaces
(integrals/two_electron/fock/jk_direct.hpp:18-96,
integrals/two_electron/fock/eri_direct_jk.hpp:17-39,
integrals/two_electron/twoe_integrals.hpp:41-113).

Same names, argument meaning and error behaviour:
  * the plan is spherical by default (the reference's only basis); pass
    ``cartesian=True`` for the Cartesian component basis (AO index =
    component index);
  * densities are M x M, finite and symmetric within 128 eps (else
    ``RuntimeError("eri_ao_tensor_hgp: density ...")``);
  * thresholds must be finite and >= 0.

``density_threshold`` (the reference's density-coupled screen) and the
heuristic ``decay_threshold`` (shell-pair decay indicator) and
``primitive_fastschwarz_threshold`` screens apply to the J/K builds.  The
FASTSCHWARZ test is made per primitive quartet; the reference tests a SIMD
packet of bra pairs together and so skips a subset of those quartets (the two
agree where every packet holds one pair).  The stream checksum endpoint takes
the Schwarz threshold only and raises NotImplementedError for the other two.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .basis import Basis

DEFAULT_SCHWARZ = 1.0e-12  # kEriPackedDefaultSchwarzThreshold (twoe_integrals.hpp:41)


@dataclass
class ScreeningOptions:
    integral_threshold: float = DEFAULT_SCHWARZ
    primitive_fastschwarz_threshold: float = 0.0
    decay_threshold: float = 0.0
    density_threshold: float = 0.0


@dataclass
class EriPackedBuildOptions:
    schwarz_threshold: float = DEFAULT_SCHWARZ
    primitive_fastschwarz_threshold: float = 0.0
    decay_threshold: float = 0.0
    density_threshold: float = 0.0


def screening_from_eri_options(o: EriPackedBuildOptions) -> ScreeningOptions:
    return ScreeningOptions(o.schwarz_threshold, o.primitive_fastschwarz_threshold, o.decay_threshold,
                            o.density_threshold)


def screening_to_eri_options(o: ScreeningOptions) -> EriPackedBuildOptions:
    return EriPackedBuildOptions(o.integral_threshold, o.primitive_fastschwarz_threshold, o.decay_threshold,
                                 o.density_threshold)


def _check_threshold(name, v):
    if not (isinstance(v, (int, float)) and math.isfinite(v) and v >= 0.0):
        raise RuntimeError(f"eri_ao_tensor_hgp: {name} must be finite and >= 0.")


def _validate_options(o: ScreeningOptions):
    _check_threshold("schwarz_threshold", o.integral_threshold)
    _check_threshold("primitive_fastschwarz_threshold", o.primitive_fastschwarz_threshold)
    _check_threshold("decay_threshold", o.decay_threshold)
    _check_threshold("density_threshold", o.density_threshold)
    if o.density_threshold > 0.0 and o.decay_threshold > 0.0:
        raise RuntimeError("eri_ao_tensor_hgp: density screening requires Schwarz bounds and cannot be combined "
                           "with the heuristic decay metric.")


@dataclass
class PlanProfile:
    """Subset of EriPackedPlanProfile (twoe_integrals.hpp:68-89)."""
    M: int = 0
    shell_groups: int = 0
    pair_groups: int = 0
    primitive_pairs: int = 0


class DirectPlan:
    """GPU-resident plan: grouped shells, primitive-pair SoA and Schwarz bounds."""

    def __init__(self, basis: Basis, options: ScreeningOptions | None = None, cartesian: bool = False,
                 device: int = 0):
        lib = _lib.load()
        _lib.check(lib.tlno_gpu_init(int(device)))
        self.basis = basis
        self.cartesian = bool(cartesian)
        self.options = options or ScreeningOptions()
        _validate_options(self.options)
        self._arrays = (np.ascontiguousarray(basis.centers, dtype=np.float64).reshape(-1),
                        np.ascontiguousarray(basis.ang, dtype=np.int32).reshape(-1),
                        np.ascontiguousarray(basis.nprim, dtype=np.int32),
                        np.ascontiguousarray(basis.exps, dtype=np.float64),
                        np.ascontiguousarray(basis.coefs, dtype=np.float64))
        c, a, n, e, f = self._arrays
        h = C.c_void_p()
        _lib.check(lib.tlno_gpu_plan_create(len(n), c.ctypes.data_as(_lib.D), a.ctypes.data_as(_lib.I),
                                            n.ctypes.data_as(_lib.I), e.ctypes.data_as(_lib.D),
                                            f.ctypes.data_as(_lib.D), 0 if cartesian else 1, C.byref(h)))
        self._h = h
        M = C.c_int()
        ng = C.c_int()
        npair = C.c_int()
        npp = C.c_longlong()
        _lib.check(lib.tlno_gpu_plan_info(h, C.byref(M), C.byref(ng), C.byref(npair), C.byref(npp)))
        self.profile = PlanProfile(M.value, ng.value, npair.value, npp.value)

    @property
    def handle(self):
        return self._h

    @property
    def M(self):
        return self.profile.M

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().tlno_gpu_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def jk(self, density, want_j=True, want_k=True, schwarz=None, density_threshold=None, fastschwarz=None,
           decay=None):
        """Raw J and K (either may be skipped) on host arrays."""
        thr = self.options.integral_threshold if schwarz is None else schwarz
        _check_threshold("schwarz_threshold", thr)
        dthr = self.options.density_threshold if density_threshold is None else density_threshold
        _check_threshold("density_threshold", dthr)
        fs = self.options.primitive_fastschwarz_threshold if fastschwarz is None else fastschwarz
        dec = self.options.decay_threshold if decay is None else decay
        _check_threshold("primitive_fastschwarz_threshold", fs)
        _check_threshold("decay_threshold", dec)
        _lib.check(_lib.load().tlno_gpu_set_density_threshold(self._h, float(dthr)))
        _lib.check(_lib.load().tlno_gpu_set_primitive_screens(self._h, float(fs), float(dec)))
        M = self.M
        D = np.ascontiguousarray(density, dtype=np.float64)
        if D.shape != (M, M):
            raise RuntimeError("eri_ao_tensor_hgp: density must be spherical M x M row-major.")
        J = np.empty((M, M)) if want_j else None
        K = np.empty((M, M)) if want_k else None
        _lib.check(_lib.load().tlno_gpu_build_jk(self._h, D.ctypes.data_as(_lib.D),
                                                 J.ctypes.data_as(_lib.D) if want_j else None,
                                                 K.ctypes.data_as(_lib.D) if want_k else None, float(thr)))
        return J, K

    def set_team_threshold(self, min_outputs):
        """FP64 classes with >= min_outputs Cartesian values per shell quartet
        use the one-wave-per-quartet evaluator (0: lane evaluator everywhere)."""
        _lib.check(_lib.load().tlno_gpu_set_team_threshold(self._h, int(min_outputs)))

    def autotune(self, schwarz=None, repeats=2):
        """Pick the faster of the lane and team evaluators per class pair on
        this plan, and whether the FP32 kernels pay for a class's FP32 share;
        returns the timings and choices."""
        import json
        thr = self.options.integral_threshold if schwarz is None else schwarz
        buf = C.create_string_buffer(8192)
        _lib.check(_lib.load().tlno_gpu_plan_autotune(self._h, float(thr), int(repeats), buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_mixed_precision(self, fp32_threshold):
        """Route shell quartets with Q_ab Q_cd max|D| < fp32_threshold to the
        FP32 recurrences (0 disables); J/K accumulation stays FP64."""
        _lib.check(_lib.load().tlno_gpu_set_mixed_precision(self._h, float(fp32_threshold)))

    def eri_checksum(self, schwarz=None):
        """Sum of the canonical unique spherical ERIs over Schwarz-surviving
        shell quartets (the reference's packed-stream checksum)."""
        thr = self.options.integral_threshold if schwarz is None else schwarz
        _check_threshold("schwarz_threshold", thr)
        v = C.c_double()
        _lib.check(_lib.load().tlno_gpu_eri_checksum(self._h, float(thr), C.byref(v)))
        return v.value

    def eri_fingerprint(self, schwarz=None):
        """The reference's indexed stream fingerprint of the same canonical
        values (its TLNO_VALIDATE_RYS_STREAM_FINGERPRINT build,
        rys_engine.ipp:87-105, 270-300): checksum, absolute_sum, square_sum,
        indexed_moment_1..3 (values weighted by splitmix64 hashes of their
        canonical quartet index), max_abs and canonical_count."""
        thr = self.options.integral_threshold if schwarz is None else schwarz
        _check_threshold("schwarz_threshold", thr)
        out = np.zeros(7)
        n = C.c_ulonglong(0)
        _lib.check(_lib.load().tlno_gpu_eri_fingerprint(self._h, float(thr), out.ctypes.data_as(_lib.D), C.byref(n)))
        keys = ("checksum", "absolute_sum", "square_sum", "indexed_moment_1", "indexed_moment_2",
                "indexed_moment_3", "max_abs")
        r = {k: float(v) for k, v in zip(keys, out)}
        r["canonical_count"] = int(n.value)
        return r

    def set_learned_screen(self, params=None, margin=0.0):
        """Turn the learned shell-quartet screen on (params: 6401 floats of a
        trained network, margin in decades) or off (params None)."""
        lib = _lib.load()
        if params is None:
            _lib.check(lib.tlno_gpu_set_learned_screen(self._h, None, 0.0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (6401,):
            raise ValueError("learned screen: expected 6401 network parameters")
        _lib.check(lib.tlno_gpu_set_learned_screen(self._h, p.ctypes.data_as(_lib.F), float(margin)))

    def last_route_stats(self):
        """Per-quartet routing counts of the last build: candidates, kept by
        the density test, kept after the learned screen, sent to FP32."""
        out = np.zeros(4, np.int64)
        _lib.check(_lib.load().tlno_gpu_last_route_stats(self._h, out.ctypes.data_as(_lib.LL)))
        return {"candidates": int(out[0]), "density_kept": int(out[1]), "kept": int(out[2]), "fp32": int(out[3])}

    def last_mixed_stats(self):
        sq = C.c_longlong()
        pq = C.c_longlong()
        md = C.c_double()
        _lib.check(_lib.load().tlno_gpu_last_mixed_stats(self._h, C.byref(sq), C.byref(pq), C.byref(md)))
        return {"fp32_shell_quartets": sq.value, "fp32_prim_quartets": pq.value, "max_abs_density": md.value}

    def last_build_stats(self):
        sq = C.c_longlong()
        pq = C.c_longlong()
        _lib.check(_lib.load().tlno_gpu_last_build_stats(self._h, C.byref(sq), C.byref(pq)))
        return sq.value, pq.value

    def eri_cart_blocks(self, quads):
        """Unit-diagonal Cartesian blocks for group quadruples [(ga, gb, gc, gd), ...]."""
        from .basis import ncart
        quads = np.ascontiguousarray(quads, dtype=np.int32).reshape(-1, 4)
        Ls = self.group_L()
        sizes = [int(np.prod([ncart(Ls[g]) for g in q])) for q in quads]
        offs = np.zeros(len(quads), dtype=np.int64)
        if len(quads):
            offs[1:] = np.cumsum(sizes)[:-1]
        out = np.zeros(int(sum(sizes)))
        _lib.check(_lib.load().tlno_gpu_eri_cart_blocks(self._h, len(quads), quads.ctypes.data_as(_lib.I),
                                                        offs.ctypes.data_as(_lib.LL), out.ctypes.data_as(_lib.D)))
        res = []
        for q, o, s in zip(quads, offs, sizes):
            res.append(out[o:o + s].reshape([ncart(Ls[g]) for g in q]))
        return res

    def pairs(self):
        """(group_a, group_b, Schwarz bound, class_begin[16]) in plan pair order (classes ss, ps, pp,
        ds, dp, dd, fs, fp, fd, ff, gs, gp, gd, gf, gg)."""
        n = self.profile.pair_groups
        ga = np.zeros(n, np.int32)
        gb = np.zeros(n, np.int32)
        Q = np.zeros(n)
        cb = np.zeros(16, np.int32)
        _lib.check(_lib.load().tlno_gpu_plan_pairs(self._h, ga.ctypes.data_as(_lib.I), gb.ctypes.data_as(_lib.I),
                                                   Q.ctypes.data_as(_lib.D), cb.ctypes.data_as(_lib.I)))
        return ga, gb, Q, cb

    def quartet_maxabs(self, P, Q):
        """max |(ab|cd)| of the unit-diagonal block of each (P[i] | Q[i]) (plan pair indices)."""
        P = np.ascontiguousarray(P, dtype=np.int32)
        Q = np.ascontiguousarray(Q, dtype=np.int32)
        out = np.zeros(len(P))
        _lib.check(_lib.load().tlno_gpu_quartet_maxabs(self._h, len(P), P.ctypes.data_as(_lib.I),
                                                       Q.ctypes.data_as(_lib.I), out.ctypes.data_as(_lib.D)))
        return out

    def group_L(self):
        """L of each shell group of the plan, in first-appearance order."""
        self._load_groups()
        return self._group_L

    def group_components(self):
        """basis index of each group's components in descending-lex order."""
        self._load_groups()
        return self._group_comps

    def _load_groups(self):
        if hasattr(self, "_group_L"):
            return
        ng = self.profile.shell_groups
        L = np.zeros(ng, np.int32)
        comps = np.zeros(15 * ng, np.int32)
        _lib.check(_lib.load().tlno_gpu_plan_groups(self._h, ng, L.ctypes.data_as(_lib.I),
                                                    comps.ctypes.data_as(_lib.I)))
        self._group_L = [int(x) for x in L]
        self._group_comps = [comps[15 * g:15 * g + (l + 1) * (l + 2) // 2].copy() for g, l in enumerate(self._group_L)]


@dataclass
class CoulombJResult:
    j: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    M: int = 0


@dataclass
class ExchangeKResult:
    k: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    M: int = 0


@dataclass
class DirectResult:
    g: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    M: int = 0


def build_direct_plan(basis: Basis, options: ScreeningOptions | None = None, cartesian=False) -> DirectPlan:
    return DirectPlan(basis, options, cartesian=cartesian)


def _opts(plan, options):
    o = options or plan.options
    _validate_options(o)
    return o


def build_coulomb_j(plan: DirectPlan, density, options: ScreeningOptions | None = None) -> CoulombJResult:
    o = _opts(plan, options)
    J, _ = plan.jk(density, True, False, o.integral_threshold, o.density_threshold, o.primitive_fastschwarz_threshold,
                   o.decay_threshold)
    return CoulombJResult(J, plan.M)


def build_exchange_k(plan: DirectPlan, density, options: ScreeningOptions | None = None) -> ExchangeKResult:
    """K[a,c] = sum_(b,d) P[b,d] (ab|cd), for a real symmetric density."""
    o = _opts(plan, options)
    _, K = plan.jk(density, False, True, o.integral_threshold, o.density_threshold, o.primitive_fastschwarz_threshold,
                   o.decay_threshold)
    return ExchangeKResult(K, plan.M)


def build_rhf_g(plan: DirectPlan, density, options: ScreeningOptions | None = None) -> DirectResult:
    """G = J - K/2 from one traversal."""
    o = _opts(plan, options)
    J, K = plan.jk(density, True, True, o.integral_threshold, o.density_threshold, o.primitive_fastschwarz_threshold,
                   o.decay_threshold)
    return DirectResult(J - 0.5 * K, plan.M)


# --- the twoe:: names (EriPackedPlan is the same GPU plan object) ---
def build_eri_packed_plan_for_G_rys(basis: Basis, options: EriPackedBuildOptions | None = None) -> DirectPlan:
    return DirectPlan(basis, screening_from_eri_options(options or EriPackedBuildOptions()))


def adopt_eri_plan_for_transition(plan: DirectPlan) -> DirectPlan:
    return plan


@dataclass
class EriPackedGData:
    """values = [checksum of the canonical spherical stream]
    (twoe_integrals.hpp:37-39, rys_engine.ipp:17096-17103); with
    fingerprint=True [checksum, indexed_moment_1..3, canonical_count], what
    the reference's fingerprint build returns (native_quartets.ipp:759-765)."""
    values: list = field(default_factory=list)


def build_eri_packed_for_G_rys(plan: DirectPlan, options: EriPackedBuildOptions | None = None,
                               fingerprint: bool = False) -> EriPackedGData:
    o = options or screening_to_eri_options(plan.options)
    _check_threshold("schwarz_threshold", o.schwarz_threshold)
    _check_threshold("decay_threshold", o.decay_threshold)
    _check_threshold("primitive_fastschwarz_threshold", o.primitive_fastschwarz_threshold)
    for name in ("primitive_fastschwarz_threshold", "decay_threshold"):
        if getattr(o, name) != 0.0:
            raise NotImplementedError(f"{name} screening is not available in the GPU engine")
    # the stream ignores density_threshold: there is no density to bound against
    if fingerprint:
        f = plan.eri_fingerprint(o.schwarz_threshold)
        return EriPackedGData([f["checksum"], f["indexed_moment_1"], f["indexed_moment_2"], f["indexed_moment_3"],
                               float(f["canonical_count"])])
    return EriPackedGData([plan.eri_checksum(o.schwarz_threshold)])


@dataclass
class EriDirectJKResult:
    g: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    M: int = 0


def build_jk_direct_for_G_rys(plan: DirectPlan, density, options: EriPackedBuildOptions | None = None):
    o = screening_from_eri_options(options or EriPackedBuildOptions())
    r = build_rhf_g(plan, density, o)
    return EriDirectJKResult(r.g, r.M)


# --- incremental SCF state (eri_direct_jk.hpp:41-135, direct_build.ipp:562-775) ---
import enum


class EriDirectJKUpdateStatus(enum.IntEnum):
    uninitialized = 0
    exact_rebuild = 1
    incremental_update = 2
    skipped_unchanged = 3


@dataclass
class EriDirectJKIncrementalOptions:
    base_options: EriPackedBuildOptions = field(default_factory=EriPackedBuildOptions)
    exact_rebuild_interval: int = 8
    delta_density_threshold: float = 0.0
    use_first_delta_density_threshold: bool = False
    first_delta_density_threshold: float = 0.0


@dataclass
class EriDirectJKIncrementalCounters:
    updates: int = 0
    exact_rebuilds: int = 0
    incremental_updates: int = 0
    skipped_updates: int = 0


class EriDirectJKIncrementalState:
    """G for SCF iterations: exact rebuilds from the full density (first call,
    forced, or every exact_rebuild_interval changed densities); in between
    G += G(dP), dP optionally density-screened.  A failed update leaves the
    state unchanged."""

    def __init__(self, plan: DirectPlan, options: EriDirectJKIncrementalOptions | None = None):
        import copy
        self.plan = plan
        self.options = copy.deepcopy(options or EriDirectJKIncrementalOptions())
        o = self.options
        if int(o.exact_rebuild_interval) < 1:
            raise RuntimeError("eri_ao_tensor_hgp: incremental exact_rebuild_interval must be >= 1.")
        b = o.base_options
        for name in ("schwarz_threshold", "primitive_fastschwarz_threshold", "decay_threshold"):
            v = getattr(b, name)
            if not (math.isfinite(v) and v >= 0.0):
                raise RuntimeError(f"eri_ao_tensor_hgp: incremental {name} must be finite and >= 0.")
        if not (math.isfinite(b.density_threshold) and b.density_threshold >= 0.0):
            raise RuntimeError("eri_ao_tensor_hgp: incremental base density_threshold must be finite and >= 0.")
        if b.schwarz_threshold != 0.0 and b.decay_threshold != 0.0:
            raise RuntimeError("eri_ao_tensor_hgp: incremental base options cannot combine Schwarz and decay "
                               "screening.")
        for name in ("delta_density_threshold", "first_delta_density_threshold"):
            v = getattr(o, name)
            if not (math.isfinite(v) and v >= 0.0):
                raise RuntimeError(f"eri_ao_tensor_hgp: incremental {name} must be finite and >= 0.")
        if ((o.delta_density_threshold > 0.0 or (o.use_first_delta_density_threshold
                                                  and o.first_delta_density_threshold > 0.0))
                and b.decay_threshold > 0.0):
            raise RuntimeError("eri_ao_tensor_hgp: incremental delta-density screening cannot be combined with "
                               "the heuristic decay metric.")
        b.density_threshold = 0.0  # exact rebuilds never screen with the full density
        self.M = plan.M
        self.reset()

    def reset(self):
        self.density = None
        self.g = None
        self.counters = EriDirectJKIncrementalCounters()
        self.last_status = EriDirectJKUpdateStatus.uninitialized
        self.last_max_abs_density_delta = 0.0
        self.last_effective_density_threshold = 0.0
        self.updates_since_exact_rebuild = 0

    @property
    def initialized(self):
        return self.density is not None

    def update(self, density, force_exact_rebuild=False) -> EriDirectJKUpdateStatus:
        D = np.asarray(density, dtype=np.float64)
        if D.shape != (self.M, self.M):
            raise RuntimeError("eri_ao_tensor_hgp: incremental density must be spherical M x M row-major.")
        if not np.all(np.isfinite(D)):
            raise RuntimeError("eri_ao_tensor_hgp: incremental density contains a non-finite value.")
        delta = D - self.density if self.initialized else D
        if not np.all(np.isfinite(delta)):
            raise RuntimeError("eri_ao_tensor_hgp: incremental density delta overflowed.")
        max_abs = float(np.abs(delta).max()) if delta.size else 0.0
        if self.initialized and max_abs == 0.0 and not force_exact_rebuild:
            self.counters.updates += 1
            self.counters.skipped_updates += 1
            self.last_status = EriDirectJKUpdateStatus.skipped_unchanged
            self.last_max_abs_density_delta = 0.0
            self.last_effective_density_threshold = 0.0
            return self.last_status
        o = self.options
        exact = (not self.initialized or force_exact_rebuild
                 or self.updates_since_exact_rebuild + 1 >= int(o.exact_rebuild_interval))
        import copy
        bo = copy.copy(o.base_options)
        eff = 0.0
        if not exact:
            eff = (o.first_delta_density_threshold
                   if o.use_first_delta_density_threshold and self.updates_since_exact_rebuild == 0
                   else o.delta_density_threshold)
            bo.density_threshold = eff
        built = build_jk_direct_for_G_rys(self.plan, D if exact else delta, bo).g
        if not np.all(np.isfinite(built)):
            raise RuntimeError("eri_ao_tensor_hgp: direct J/K build produced a non-finite value.")
        g = built if exact else self.g + built
        self.density, self.g = D.copy(), g
        self.counters.updates += 1
        self.last_max_abs_density_delta = max_abs
        self.last_effective_density_threshold = eff
        if exact:
            self.counters.exact_rebuilds += 1
            self.updates_since_exact_rebuild = 0
            self.last_status = EriDirectJKUpdateStatus.exact_rebuild
        else:
            self.counters.incremental_updates += 1
            self.updates_since_exact_rebuild += 1
            self.last_status = EriDirectJKUpdateStatus.incremental_update
        return self.last_status
