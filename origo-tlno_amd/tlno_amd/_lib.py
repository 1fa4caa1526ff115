"""Loads the HIP library (libtlno_amd.so) and declares its C ABI (include/tlno_gpu.h).

There is no fallback: if the library is missing or no HIP device is usable,
every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# TLNO_AMD_LIB selects another build of the same library (e.g. a variant
# under test); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("TLNO_AMD_LIB") or os.path.join(_HERE, "libtlno_amd.so")

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
LL = C.POINTER(C.c_longlong)
F = C.POINTER(C.c_float)

# every symbol the header declares, with (restype, argtypes)
SIGNATURES = {
    "tlno_gpu_last_error": (C.c_char_p, []),
    "tlno_gpu_init": (C.c_int, [C.c_int]),
    "tlno_gpu_plan_create": (C.c_int, [C.c_int, D, I, I, D, D, C.c_int, C.POINTER(C.c_void_p)]),
    "tlno_gpu_plan_destroy": (None, [C.c_void_p]),
    "tlno_gpu_plan_info": (C.c_int, [C.c_void_p, I, I, I, LL]),
    "tlno_gpu_plan_groups": (C.c_int, [C.c_void_p, C.c_int, I, I]),
    "tlno_gpu_plan_pairs": (C.c_int, [C.c_void_p, I, I, D, I]),
    "tlno_gpu_quartet_maxabs": (C.c_int, [C.c_void_p, C.c_int, I, I, D]),
    "tlno_gpu_build_jk": (C.c_int, [C.c_void_p, D, D, D, C.c_double]),
    "tlno_gpu_build_jk_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "tlno_gpu_build_jk_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    "tlno_gpu_last_build_stats": (C.c_int, [C.c_void_p, LL, LL]),
    "tlno_gpu_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "tlno_gpu_set_mixed_precision": (C.c_int, [C.c_void_p, C.c_double]),
    "tlno_gpu_set_team_threshold": (C.c_int, [C.c_void_p, C.c_int]),
    "tlno_gpu_plan_autotune": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_char_p, C.c_int]),
    "tlno_gpu_set_density_threshold": (C.c_int, [C.c_void_p, C.c_double]),
    "tlno_gpu_set_primitive_screens": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "tlno_gpu_last_mixed_stats": (C.c_int, [C.c_void_p, LL, LL, D]),
    "tlno_gpu_class_stats": (C.c_int, [C.c_void_p, LL, LL, D, LL]),
    "tlno_gpu_class_stats_all": (C.c_int, [C.c_void_p, LL, LL, D, LL]),
    "tlno_gpu_eri_cart_blocks": (C.c_int, [C.c_void_p, C.c_int, I, LL, D]),
    "tlno_gpu_boys": (C.c_int, [C.c_int, D, C.c_int, D]),
    "tlno_gpu_rys": (C.c_int, [C.c_int, C.c_int, D, D, D]),
    "tlno_gpu_onee": (C.c_int, [C.c_void_p, C.c_int, D, D, D, D, D]),
    "tlno_gpu_onee_screened": (C.c_int, [C.c_void_p, C.c_int, D, D, C.c_double, C.c_double, D, D, D, D]),
    "tlno_gpu_eri_checksum": (C.c_int, [C.c_void_p, C.c_double, D]),
    "tlno_gpu_eri_fingerprint": (C.c_int, [C.c_void_p, C.c_double, D, C.POINTER(C.c_ulonglong)]),
    "tlno_gpu_batch_create": (C.c_int, [C.c_int, I, D, I, I, D, D, C.POINTER(C.c_void_p)]),
    "tlno_gpu_batch_destroy": (None, [C.c_void_p]),
    "tlno_gpu_batch_info": (C.c_int, [C.c_void_p, I, LL, I, LL]),
    "tlno_gpu_batch_build_jk": (C.c_int, [C.c_void_p, D, D, D, C.c_double]),
    "tlno_gpu_batch_build_jk_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                                 C.c_int]),
    "tlno_gpu_batch_autotune": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_char_p, C.c_int]),
    "tlno_gpu_batch_set_shape": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "tlno_gpu_batch_last_stats": (C.c_int, [C.c_void_p, LL, LL, D]),
    "tlno_gpu_set_learned_screen": (C.c_int, [C.c_void_p, F, C.c_double]),
    "tlno_gpu_quartet_features": (C.c_int, [C.c_void_p, C.c_int, I, I, F]),
    "tlno_gpu_last_route_stats": (C.c_int, [C.c_void_p, LL]),
    "tlno_gpu_mlp_create": (C.c_int, [C.c_int, C.c_uint, C.POINTER(C.c_void_p)]),
    "tlno_gpu_mlp_destroy": (None, [C.c_void_p]),
    "tlno_gpu_mlp_get_params": (C.c_int, [C.c_void_p, F]),
    "tlno_gpu_mlp_set_params": (C.c_int, [C.c_void_p, F]),
    "tlno_gpu_mlp_set_normalization": (C.c_int, [C.c_void_p, C.c_longlong, F]),
    "tlno_gpu_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_longlong, F, F, C.c_float, F, D]),
    "tlno_gpu_mlp_adam_step": (C.c_int, [C.c_void_p, C.c_longlong, F, F, C.c_float, C.c_float]),
    "tlno_gpu_mlp_train": (C.c_int, [C.c_void_p, C.c_longlong, F, F, C.c_int, C.c_int, C.c_float, C.c_float,
                                      C.c_uint, F]),
    "tlno_gpu_mlp_predict": (C.c_int, [C.c_void_p, C.c_longlong, F, F, C.c_int]),
}

_lib = None


class TlnoGpuError(RuntimeError):
    pass


def load():
    """Return the loaded library (symbols declared); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TlnoGpuError(
            f"HIP library not built: {LIB_PATH} (run __graft_entry__.build() or make -C origo-tlno_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise TlnoGpuError(load().tlno_gpu_last_error().decode())
