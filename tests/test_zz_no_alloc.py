"""GPU: the allocation-free J/K build (tlno_gpu_build_jk_device_noalloc)
through the C ABI, against the allocating entry point
(tlno_gpu_build_jk_device) on the same plan and density: two calls in a row,
the second on a smaller molecule, then with the density screen and the FP32
gate, and the errors for an unprepared threshold and a plan above the
scratch limit.  Tolerance: J/K max-abs 1e-12 (same kernels; FP64 atomics add
in a run-dependent order)."""
import ctypes as C

import numpy as np
import pytest

from tlno_amd import _lib, configs, hipmem
from tlno_amd.fock import DirectPlan, ScreeningOptions

pytestmark = pytest.mark.gpu

THR = 1e-12


def _run(entry, plan, dD, thr=THR):
    lib = _lib.load()
    M = plan.M
    dJ = hipmem.DeviceArray((M, M))
    dK = hipmem.DeviceArray((M, M))
    rc = getattr(lib, entry)(plan.handle, dD.ptr, dJ.ptr, dK.ptr, C.c_double(thr))
    _lib.check(rc)
    return dJ.to_host(), dK.to_host(), plan.last_build_stats()


def _same(a, b):
    Ja, Ka, sa = a
    Jb, Kb, sb = b
    assert sa == sb
    assert np.abs(Ja - Jb).max() <= 1e-12 and np.abs(Ka - Kb).max() <= 1e-12


def test_noalloc_matches_allocating_entry_point():
    big_b, big_D = configs.organic30()          # 340 Cartesian functions
    small_b, small_D = configs.water_cluster(4)  # 100 Cartesian functions
    big = DirectPlan(big_b, ScreeningOptions(integral_threshold=THR), cartesian=True)
    small = DirectPlan(small_b, ScreeningOptions(integral_threshold=THR), cartesian=True)
    dbig = hipmem.DeviceArray.from_host(big_D)
    dsmall = hipmem.DeviceArray.from_host(small_D)
    # the allocating entry point prepares each plan (work lists, streams)
    ref_big = _run("tlno_gpu_build_jk_device", big, dbig)
    ref_small = _run("tlno_gpu_build_jk_device", small, dsmall)
    # two allocation-free calls in a row, the second smaller: the scratch
    # keeps the first call's contents, which must not leak into the second
    _same(_run("tlno_gpu_build_jk_device_noalloc", big, dbig), ref_big)
    _same(_run("tlno_gpu_build_jk_device_noalloc", small, dsmall), ref_small)
    # again in the other order
    _same(_run("tlno_gpu_build_jk_device_noalloc", big, dbig), ref_big)
    # density screen + FP32 gate (routed build: per-quartet routing buffers)
    small.set_mixed_precision(1e-6)
    _lib.check(_lib.load().tlno_gpu_set_density_threshold(small.handle, 2e-9))
    ref_routed = _run("tlno_gpu_build_jk_device", small, dsmall)
    got_routed = _run("tlno_gpu_build_jk_device_noalloc", small, dsmall)
    _same(got_routed, ref_routed)
    assert small.last_route_stats()["candidates"] > 0
    assert small.last_mixed_stats()["fp32_shell_quartets"] > 0


def test_noalloc_rejects_unprepared_threshold_and_large_plans():
    b, D = configs.water_cluster(4)
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=THR), cartesian=True)
    dD = hipmem.DeviceArray.from_host(D)
    _run("tlno_gpu_build_jk_device", plan, dD)
    with pytest.raises(_lib.TlnoGpuError, match="work lists for this Schwarz threshold not prepared"):
        _run("tlno_gpu_build_jk_device_noalloc", plan, dD, thr=1e-10)
    # the same threshold still works after the refusal
    _run("tlno_gpu_build_jk_device_noalloc", plan, dD)
    bb, Db = configs.water_cluster(42)  # 1050 Cartesian functions
    big = DirectPlan(bb, ScreeningOptions(integral_threshold=THR), cartesian=True)
    assert big.M > 1024
    dDb = hipmem.DeviceArray.from_host(Db)
    with pytest.raises(_lib.TlnoGpuError, match="exceed the scratch limit of 1024"):
        _run("tlno_gpu_build_jk_device_noalloc", big, dDb)
