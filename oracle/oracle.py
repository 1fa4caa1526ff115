"""ORACLE — test infrastructure only.

ctypes bindings for
  * oracle/_build/liboracle.so      the plain-C restatement of the reference
                                    path (the checker);
  * oracle/_ref/libtlno_ref_<isa>.so, libtlno_ref_audit_<isa>.so
                                    the reference engine compiled from its own
                                    sources by oracle/ref.mk (optional; only
                                    present where it was built).

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  The product (origo-tlno_amd) never does.

Basis arguments use the reference's component layout (one entry per
Cartesian component, see core/basis/types.hpp:31-36): an object with numpy
attributes ``centers (n,3)``, ``ang (n,3) int32``, ``nprim (n,) int32``,
``exps``, ``coefs`` (concatenated primitives).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def _basis_args(b):
    centers = np.ascontiguousarray(b.centers, dtype=np.float64).reshape(-1)
    ang = np.ascontiguousarray(b.ang, dtype=np.int32).reshape(-1)
    nprim = np.ascontiguousarray(b.nprim, dtype=np.int32)
    exps = np.ascontiguousarray(b.exps, dtype=np.float64)
    coefs = np.ascontiguousarray(b.coefs, dtype=np.float64)
    keep = (centers, ang, nprim, exps, coefs)
    return keep, (len(nprim), _dp(centers), _ip(ang), _ip(nprim), _dp(exps), _dp(coefs))


# --------------------------------------------------------------------------
# the C restatement
# --------------------------------------------------------------------------
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_boys.restype = C.c_double
        L.orc_boys.argtypes = [C.c_int, C.c_double]
        L.orc_rys.argtypes = [C.c_int, C.c_double, _D, _D]
        L.orc_rys_fast.argtypes = [C.c_int, C.c_double, _D, _D]
        L.orc_rys_table_init.argtypes = [C.c_int]
        L.orc_normalize.argtypes = [C.c_int, _D, _D, _I, _D]
        L.orc_cart2sph.argtypes = [C.c_int, _D]
        L.orc_basis_create.restype = C.c_void_p
        L.orc_basis_create.argtypes = [C.c_int, _D, _I, _I, _D, _D]
        L.orc_basis_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        for f in ("orc_ngroups", "orc_ncart", "orc_nsph"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_group_info.argtypes = [C.c_void_p, C.c_int, _I, _I, _I]
        L.orc_eri_cart_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _D]
        L.orc_jk.argtypes = [C.c_void_p, C.c_int, C.c_double, _D, _D, _D, C.c_int]
        L.orc_jk_screened.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _D, _D, _D, C.c_int]
        L.orc_schwarz.argtypes = [C.c_void_p, C.c_int, _D]
        L.orc_onee.argtypes = [C.c_void_p, C.c_int, C.c_int, _D, _D, _D, _D, _D]
        _lib = L
    return _lib


def boys(m, x):
    return lib().orc_boys(int(m), float(x))


def rys(N, T, fast=False):
    r = np.zeros(N)
    w = np.zeros(N)
    if fast:
        lib().orc_rys_table_init(N)
        lib().orc_rys_fast(N, float(T), _dp(r), _dp(w))
    else:
        rc = lib().orc_rys(N, float(T), _dp(r), _dp(w))
        if rc:
            raise ValueError("orc_rys failed")
    return r, w


def normalize(exps, coefs, ang):
    e = np.ascontiguousarray(exps, dtype=np.float64)
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    a = np.ascontiguousarray(ang, dtype=np.int32)
    out = np.zeros_like(e)
    lib().orc_normalize(len(e), _dp(e), _dp(c), _ip(a), _dp(out))
    return out


def cart2sph(L):
    nc = (L + 1) * (L + 2) // 2
    out = np.zeros((2 * L + 1, nc))
    lib().orc_cart2sph(L, _dp(out))
    return out


class OracleBasis:
    """Grouped basis (shells) exactly as the reference forms them."""

    def __init__(self, basis):
        keep, args = _basis_args(basis)
        self._keep = keep
        self.h = lib().orc_basis_create(*args)
        if not self.h:
            raise ValueError(lib().orc_last_error().decode())
        self.ngroups = lib().orc_ngroups(self.h)
        self.ncart = lib().orc_ncart(self.h)
        self.nsph = lib().orc_nsph(self.h)
        self.groups = []
        for g in range(self.ngroups):
            L = np.zeros(1, np.int32)
            comps = np.zeros(15, np.int32)
            sb = np.zeros(1, np.int32)
            lib().orc_group_info(self.h, g, _ip(L), _ip(comps), _ip(sb))
            nc = (int(L[0]) + 1) * (int(L[0]) + 2) // 2
            self.groups.append((int(L[0]), comps[:nc].copy(), int(sb[0])))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_basis_destroy(self.h)
            self.h = None

    def cart_block(self, ga, gb, gc, gd):
        dims = [len(self.groups[g][1]) for g in (ga, gb, gc, gd)]
        out = np.zeros(int(np.prod(dims)))
        rc = lib().orc_eri_cart_block(self.h, ga, gb, gc, gd, _dp(out))
        if rc:
            raise ValueError(lib().orc_last_error().decode())
        return out.reshape(dims)

    def jk(self, D, spherical=False, schwarz=0.0, want_j=True, want_k=True, nthreads=0, decay=0.0,
           fastschwarz=0.0):
        n = self.nsph if spherical else self.ncart
        D = np.ascontiguousarray(D, dtype=np.float64)
        assert D.shape == (n, n)
        J = np.zeros((n, n)) if want_j else None
        K = np.zeros((n, n)) if want_k else None
        lib().orc_jk_screened(self.h, int(spherical), float(schwarz), float(decay), float(fastschwarz), _dp(D),
                              _dp(J) if want_j else None, _dp(K) if want_k else None, int(nthreads))
        return J, K

    def schwarz(self, spherical=False):
        ng = self.ngroups
        out = np.zeros(ng * (ng + 1) // 2)
        lib().orc_schwarz(self.h, int(spherical), _dp(out))
        return out

    def onee(self, Z, R, spherical=False):
        n = self.nsph if spherical else self.ncart
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
        S = np.zeros((n, n))
        T = np.zeros((n, n))
        V = np.zeros((n, n))
        lib().orc_onee(self.h, int(spherical), len(Z), _dp(Z), _dp(R), _dp(S), _dp(T), _dp(V))
        return S, T, V


# --------------------------------------------------------------------------
# stream fingerprint of a canonical ERI stream
# --------------------------------------------------------------------------
# Restates the reference's indexed fingerprint (rys_engine.ipp:87-105 for the
# weights, 270-300 for the accumulation): splitmix64's output function of
# state v (golden-gamma increment, then two xor-shift-multiply rounds and a
# final xor-shift), bits 40..63 centred and scaled to [-1, 1), three seeds,
# long-double sums.
FP_SEEDS = (0x243f6a8885a308d3, 0x13198a2e03707344, 0xa4093822299f31d0)
_U = np.uint64


def fp_mix(v):
    v = np.asarray(v, dtype=_U)
    with np.errstate(over="ignore"):
        v = v + _U(0x9e3779b97f4a7c15)
        v = (v ^ (v >> _U(30))) * _U(0xbf58476d1ce4e5b9)
        v = (v ^ (v >> _U(27))) * _U(0x94d049bb133111eb)
    return v ^ (v >> _U(31))


def fp_weight(idx, seed):
    m = (fp_mix(np.asarray(idx, dtype=_U) ^ _U(seed)) >> _U(40)) & _U(0xffffff)
    return (m.astype(np.int64) - 0x800000).astype(np.longdouble) / np.longdouble(8388608)


def canonical_index(i, j):
    """hi (hi + 1) / 2 + lo (twoe_integrals.hpp:17-34), elementwise."""
    i = np.asarray(i, dtype=_U)
    j = np.asarray(j, dtype=_U)
    hi, lo = np.maximum(i, j), np.minimum(i, j)
    return hi * (hi + _U(1)) // _U(2) + lo


def stream_fingerprint(lhs_pair, rhs_pair, values):
    """Fingerprint of a canonical stream given as (lhs pair, rhs pair, value)."""
    x = np.asarray(values, dtype=np.longdouble)
    idx = canonical_index(lhs_pair, rhs_pair)
    out = {"checksum": x.sum(), "absolute_sum": np.abs(x).sum(), "square_sum": (x * x).sum(),
           "max_abs": np.abs(x).max() if len(x) else np.longdouble(0), "canonical_count": len(x)}
    for k, seed in enumerate(FP_SEEDS):
        out[f"indexed_moment_{k + 1}"] = (x * fp_weight(idx, seed)).sum()
    return {k: (v if k == "canonical_count" else float(v)) for k, v in out.items()}


# --------------------------------------------------------------------------
# the compiled reference (oracle/_ref), optional
# --------------------------------------------------------------------------
def host_isa():
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
    except OSError:
        return "avx2"
    return "avx512" if " avx512f" in flags else "avx2"


_ref = {}


def ref_lib(kind="api"):
    """kind 'api' -> libtlno_ref_<isa>.so, 'audit' -> libtlno_ref_audit_<isa>.so"""
    if kind in _ref:
        return _ref[kind]
    isa = host_isa()
    name = {"api": f"libtlno_ref_{isa}.so", "audit": f"libtlno_ref_audit_{isa}.so",
            "fp": "libtlno_ref_fp_avx512.so"}[kind]
    path = os.path.join(_HERE, "_ref", name)
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    if kind == "api":
        L.ref_last_error.restype = C.c_char_p
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_normalize.argtypes = [C.c_int, _D, _D, _I, _D]
        L.ref_boys_single.restype = C.c_double
        L.ref_boys_single.argtypes = [C.c_int, C.c_double]
        L.ref_boys_r000_all.argtypes = [C.c_int, C.c_double, C.c_double, _D]
        L.ref_plan_M.argtypes = [C.c_int, _D, _I, _I, _D, _D]
        L.ref_direct_build.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_double, C.c_double,
                                       C.c_int, _D, C.c_int, _D, C.c_int,
                                       C.POINTER(C.c_longlong)]
        L.ref_direct_build_ex.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_int, _D, C.c_int, _D, C.c_int, C.POINTER(C.c_longlong)]
        L.ref_eri_checksum.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_double, _D]
        L.ref_onee.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_int, _I, _D, C.c_int, C.c_int,
                               _D, _D, _D, _I]
    elif kind == "fp":
        L.ref_fp_last_error.restype = C.c_char_p
        L.ref_fp_stream.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_double, _D]
    else:
        L.ref_audit_last_error.restype = C.c_char_p
        L.ref_audit_groups.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_int, _I, _I, _I]
        L.ref_audit_cart_block.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_int, C.c_int,
                                           C.c_int, C.c_int, _D]
        L.ref_audit_rys.argtypes = [C.c_int, C.c_double, C.c_int, _D, _D]
        L.ref_audit_boys_hgp.restype = C.c_double
        L.ref_audit_boys_hgp.argtypes = [C.c_int, C.c_double]
        L.ref_audit_sph_stream.restype = C.c_longlong
        L.ref_audit_sph_stream.argtypes = [C.c_int, _D, _I, _I, _D, _D, C.c_double]
        L.ref_audit_stream_fetch.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_uint), _D]
    _ref[kind] = L
    return L


class Reference:
    """The compiled reference engine (oracle/_ref).  Raises if it is absent."""

    def __init__(self):
        self.api = ref_lib("api")
        self.audit = ref_lib("audit")
        if self.api is None:
            raise RuntimeError("compiled reference (oracle/_ref) is not available")

    def err(self):
        return self.api.ref_last_error().decode()

    def normalize(self, exps, coefs, ang):
        e = np.ascontiguousarray(exps, dtype=np.float64)
        c = np.ascontiguousarray(coefs, dtype=np.float64)
        a = np.ascontiguousarray(ang, dtype=np.int32)
        out = np.zeros_like(e)
        if self.api.ref_normalize(len(e), _dp(e), _dp(c), _ip(a), _dp(out)):
            raise ValueError(self.err())
        return out

    def boys(self, m, x):
        return self.api.ref_boys_single(int(m), float(x))

    def plan_M(self, basis):
        keep, args = _basis_args(basis)
        M = self.api.ref_plan_M(*args)
        if M < 0:
            raise ValueError(self.err())
        return M

    def direct(self, basis, D, which, schwarz=0.0, density_threshold=0.0, threads=1,
               repeats=1, fastschwarz=0.0, decay=0.0):
        """which: 'J', 'K' or 'G'. Spherical M x M density. Returns (matrix, seconds)."""
        keep, args = _basis_args(basis)
        M = D.shape[0]
        D = np.ascontiguousarray(D, dtype=np.float64)
        out = np.zeros((M, M))
        ns = C.c_longlong(0)
        self.api.ref_set_threads(int(threads))
        rc = self.api.ref_direct_build_ex(*args, float(schwarz), float(density_threshold), float(fastschwarz),
                                          float(decay), {"J": 0, "K": 1, "G": 2}[which], _dp(D), M, _dp(out),
                                          int(repeats), C.byref(ns))
        if rc:
            raise ValueError(self.err())
        return out, ns.value * 1e-9 / max(1, repeats)

    def checksum(self, basis, schwarz=0.0):
        keep, args = _basis_args(basis)
        v = np.zeros(1)
        self.api.ref_set_threads(1)
        if self.api.ref_eri_checksum(*args, float(schwarz), _dp(v)):
            raise ValueError(self.err())
        return float(v[0])

    def fingerprint(self, basis, schwarz=0.0):
        """The reference's stream fingerprint {checksum, m1, m2, m3, count}
        (its TLNO_VALIDATE_RYS_STREAM_FINGERPRINT build, oracle/ref_fp_shim.cpp)."""
        L = ref_lib("fp")
        if L is None:
            raise RuntimeError("compiled reference fingerprint build (oracle/_ref) is not available")
        keep, args = _basis_args(basis)
        v = np.zeros(5)
        self.api.ref_set_threads(1)  # the packed endpoint runs on one thread
        if L.ref_fp_stream(*args, float(schwarz), _dp(v)):
            raise ValueError(L.ref_fp_last_error().decode())
        return v

    def onee(self, basis, Z, R, spherical=False, screening=True):
        keep, args = _basis_args(basis)
        Zi = np.ascontiguousarray(Z, dtype=np.int32)
        Rr = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
        n = self.plan_M(basis) if spherical else len(basis.nprim)
        S = np.zeros((n, n))
        T = np.zeros((n, n))
        V = np.zeros((n, n))
        nout = np.zeros(1, np.int32)
        if self.api.ref_onee(*args, len(Zi), _ip(Zi), _dp(Rr), int(spherical), int(screening),
                             _dp(S), _dp(T), _dp(V), _ip(nout)):
            raise ValueError(self.err())
        assert nout[0] == n
        return S, T, V

    # ---- audit TU ----
    def groups(self, basis):
        keep, args = _basis_args(basis)
        ng = np.zeros(1, np.int32)
        L = np.zeros(4096, np.int32)
        comps = np.zeros(4096 * 15, np.int32)
        if self.audit.ref_audit_groups(*args, 4096, _ip(ng), _ip(L), _ip(comps)):
            raise ValueError(self.audit.ref_audit_last_error().decode())
        out = []
        for g in range(int(ng[0])):
            nc = (int(L[g]) + 1) * (int(L[g]) + 2) // 2
            out.append((int(L[g]), comps[g * 15:g * 15 + nc].copy()))
        return out

    def cart_block(self, basis, ga, gb, gc, gd, groups=None):
        groups = groups or self.groups(basis)
        dims = [len(groups[g][1]) for g in (ga, gb, gc, gd)]
        out = np.zeros(int(np.prod(dims)))
        keep, args = _basis_args(basis)
        if self.audit.ref_audit_cart_block(*args, ga, gb, gc, gd, _dp(out)):
            raise ValueError(self.audit.ref_audit_last_error().decode())
        return out.reshape(dims)

    def rys(self, N, T, mode=0):
        r = np.zeros(N)
        w = np.zeros(N)
        self.audit.ref_audit_rys(int(N), float(T), int(mode), _dp(r), _dp(w))
        return r, w

    def boys_hgp(self, m, x):
        return self.audit.ref_audit_boys_hgp(int(m), float(x))

    def sph_stream(self, basis, schwarz=0.0):
        keep, args = _basis_args(basis)
        n = self.audit.ref_audit_sph_stream(*args, float(schwarz))
        if n < 0:
            raise ValueError(self.audit.ref_audit_last_error().decode())
        lhs = np.zeros(n, np.uint32)
        rhs = np.zeros(n, np.uint32)
        val = np.zeros(n)
        self.audit.ref_audit_stream_fetch(lhs.ctypes.data_as(C.POINTER(C.c_uint)),
                                          rhs.ctypes.data_as(C.POINTER(C.c_uint)), _dp(val))
        return lhs, rhs, val
