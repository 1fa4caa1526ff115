"""MI355X (gfx950) Gaussian-integral engine: Boys, Rys ERIs, S/T/V and the
direct J/K Fock build, behind the reference's one- and two-electron API.

Host code mirrors the reference's interfaces; every computation runs in the
HIP library ``libtlno_amd.so`` (C ABI in include/tlno_gpu.h)."""
from . import basis, configs  # noqa: F401
from .basis import Basis  # noqa: F401

__all__ = ["basis", "configs", "Basis"]
