"""ctypes binding of the reference-compatible one-electron C ABI
(include/tlno_onee.h, exported by libtlno_amd_cpp.so) for the tests, and the
inputs of the golden fixture tests/golden/onee_capi.npz: the reference's own
test molecules and the basis parameters its tlno_onee_basis_json exported."""
import ctypes as C
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "origo-tlno_amd", "tlno_amd", "libtlno_amd_cpp.so")
GOLD = os.path.join(ROOT, "tests", "golden", "onee_capi.npz")
CASES = (("h2o", "sto-3g"), ("gly", "def2-svp"))
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        V = C.c_void_p
        L.tlno_onee_create_from_inp.restype = V
        L.tlno_onee_create_from_inp.argtypes = [C.c_char_p, C.c_char_p]
        L.tlno_onee_create.restype = V
        L.tlno_onee_basis_json.restype = C.c_char_p
        L.tlno_onee_basis_json.argtypes = [C.c_char_p, C.c_char_p]
        for f in ("nbf", "nbf_sph"):
            getattr(L, f"tlno_onee_{f}").argtypes = [V]
        for f in ("prepare_plan", "eval_full", "eval_cached", "eval_full_sph", "eval_cached_sph", "destroy", "noop"):
            getattr(L, f"tlno_onee_{f}").argtypes = [V]
        D = C.POINTER(C.c_double)
        L.tlno_onee_copy.argtypes = [V, D, D, D]
        L.tlno_onee_copy_sph.argtypes = [V, D, D, D]
        L.tlno_onee_pair_stats.argtypes = [V] + [C.POINTER(C.c_int)] * 3
        L.tlno_onee_nuclear_pair_stats.argtypes = [V, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.tlno_onee_set_screening.argtypes = [C.c_int]
        L.tlno_onee_set_group_pair_decay_eps.argtypes = [C.c_double]
        L.tlno_onee_group_pair_decay_eps.restype = C.c_double
        L.tlno_onee_set_nuclear_screen_eps.argtypes = [C.c_double]
        L.tlno_onee_nuclear_screen_eps.restype = C.c_double
        _lib = L
    return _lib


def gold():
    return np.load(GOLD)


def write_inputs(tmpdir, mol, basis):
    """The molecule's .inp text and a JSON basis file of the elements the
    reference exported for `basis`; returns both paths."""
    g = gold()
    inp = os.path.join(tmpdir, f"{mol}.inp")
    with open(inp, "w") as f:
        f.write(str(g[f"{mol}_inp"]))
    prefix = f"{basis}_"
    elements = {k[len(prefix):-5]: json.loads(str(g[k])) for k in g.files if k.startswith(prefix) and k.endswith("_json")}
    bas = os.path.join(tmpdir, f"{basis}.json")
    with open(bas, "w") as f:
        json.dump(elements, f)
    return inp, bas


def evaluate(h, n, ns, cached=True):
    L = lib()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    out = {}
    if cached:
        L.tlno_onee_prepare_plan(h)
        L.tlno_onee_eval_cached(h)
    else:
        L.tlno_onee_eval_full(h)
    S, T, V = (np.zeros((n, n)) for _ in range(3))
    L.tlno_onee_copy(h, dp(S), dp(T), dp(V))
    out.update(S=S, T=T, V=V)
    if cached:
        L.tlno_onee_eval_cached_sph(h)
    else:
        L.tlno_onee_eval_full_sph(h)
    Ss, Ts, Vs = (np.zeros((ns, ns)) for _ in range(3))
    L.tlno_onee_copy_sph(h, dp(Ss), dp(Ts), dp(Vs))
    out.update(Ssph=Ss, Tsph=Ts, Vsph=Vs)
    g, tot, kept = C.c_int(), C.c_int(), C.c_int()
    L.tlno_onee_pair_stats(h, C.byref(g), C.byref(tot), C.byref(kept))
    nt, nk = C.c_longlong(), C.c_longlong()
    L.tlno_onee_nuclear_pair_stats(h, C.byref(nt), C.byref(nk))
    out["pair_stats"] = np.array([g.value, tot.value, kept.value])
    out["nuclear_pair_stats"] = np.array([nt.value, nk.value])
    return out
