#!/usr/bin/env python3
"""Benchmark: direct J/K Fock builds per second on one MI355X.

Workload (default): organic30 — 30 atoms uniform in a 12 A cube (minimum
separation 1 A, C/N/O/H .3/.1/.1/.5), heavy atoms 3s2p1d and H 2s1p
Cartesian shells with 1-6 primitives (exponents log-uniform in [0.08, 5000]),
340 Cartesian functions; density D = C C^T (C 340 x 60, N(0,1), ||D||_F = 1);
Schwarz threshold 1e-12; FP64 throughout.

A step is one J+K build from a density already resident in HBM (one ERI
traversal producing both matrices), through the C ABI's device entry point.
The plan (shell pairs, Schwarz bounds, work lists) is built before timing, as
in the reference (plan and build are separate API calls).

Reported besides the headline value:
  * ERI throughput: shell quartets/s, primitive quartets/s, Cartesian
    integrals/s after screening;
  * roofline of the dominant class kernel (bra/ket class pair with the most
    time): algorithmic flops per launch (tlno_amd.flops model) over its
    average HIP-event-timed duration on the plan's stream, against the
    MI355X FP64 peak (78.6 TFLOP/s; vector and matrix FP64 peaks are equal);
  * cpu_baseline: the reference engine built from its own sources
    (oracle/_ref) doing the same build (G = J - K/2, spherical d, its only
    basis) on this host, one thread — the reference's K and G endpoints
    refuse more than one OpenMP thread.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "origo-tlno_amd"), ROOT]

import numpy as np  # noqa: E402

from tlno_amd import _lib, configs, flops, hipmem  # noqa: E402
from tlno_amd.fock import DirectPlan, ScreeningOptions  # noqa: E402

FP64_PEAK_TFLOPS = 78.6


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel as measured by the FETCH_SIZE and
    WRITE_SIZE counter runs recorded under profiles/ (None when not measured)."""
    try:
        with open(os.path.join(ROOT, "profiles", "jk_organic30_traffic.json")) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return rec["bytes_per_launch"] if rec else None


def gpu_run(b, D, steps, warmup, thr):
    lib = _lib.load()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=thr), cartesian=True)
    # the project's autotuner picks the lane or team evaluator per class pair
    t0 = time.perf_counter()
    plan.tuning = plan.autotune(schwarz=thr, repeats=2)
    plan.tuning_seconds = time.perf_counter() - t0
    M = plan.M
    dD = hipmem.DeviceArray.from_host(D)
    dJ = hipmem.DeviceArray((M, M))
    dK = hipmem.DeviceArray((M, M))

    def one():
        _lib.check(lib.tlno_gpu_build_jk_device(plan.handle, dD.ptr, dJ.ptr, dK.ptr, thr))

    for _ in range(warmup):
        one()
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    hipmem.synchronize()
    wall = time.perf_counter() - t0
    # per-class kernel timing (separate pass, so event records do not sit in the timed loop)
    _lib.check(lib.tlno_gpu_set_timing(plan.handle, 1))
    for _ in range(max(3, min(steps, 10))):
        one()
    sq = (C.c_longlong * 36)()
    pq = (C.c_longlong * 36)()
    ms = (C.c_double * 36)()
    nb = C.c_longlong()
    _lib.check(lib.tlno_gpu_class_stats(plan.handle, sq, pq, ms, C.byref(nb)))
    _lib.check(lib.tlno_gpu_set_timing(plan.handle, 0))
    J = dJ.to_host()
    K = dK.to_host()
    classes = []
    for X in range(6):
        for Y in range(X + 1):
            i = 6 * X + Y
            if sq[i] == 0:
                continue
            fl = flops.class_flops(X, Y, pq[i], sq[i])
            t = ms[i] / nb.value * 1e-3
            classes.append({"class": f"({flops.CLASS_NAME[X]}|{flops.CLASS_NAME[Y]})", "shell_quartets": sq[i],
                            "prim_quartets": pq[i], "ms": ms[i] / nb.value, "gflop": fl / 1e9,
                            "tflops": fl / t / 1e12 if t > 0 else 0.0,
                            "integrals": sq[i] * flops.n_out(*flops.CLASS_L[X], *flops.CLASS_L[Y])})
    return plan, wall, classes, J, K


SCREEN_TAU = 2e-9  # density threshold of the screened water40 builds (J/K error budget 1e-7)


def mixed_run(steps=3, gate=1e-6, tau=SCREEN_TAU):
    """water40 (1000 functions), densities resident in HBM, Schwarz 1e-12:
      fp64           the plain FP64 GPU build;
      mixed          FP32 recurrences for Q_ab Q_cd max|D| < gate;
      density+mixed  plus the reference's density-coupled screen at tau (the
                     method without the network);
      learned+mixed  plus the learned screen (shipped weights): the network's
                     estimate replaces Q_ab Q_cd in the screen and in the gate.
    Errors are max-abs / RMS against the FP64 build."""
    from tlno_amd import learned
    lib = _lib.load()
    b, D = configs.water40()
    thr = 1e-12
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=thr), cartesian=True)
    tuning = plan.autotune(schwarz=thr, repeats=1)
    M = plan.M
    dD = hipmem.DeviceArray.from_host(D)
    dJ = hipmem.DeviceArray((M, M))
    dK = hipmem.DeviceArray((M, M))
    params, margin = learned.load_weights()
    runs = {}
    for name, g, t, use_net in (("fp64", 0.0, 0.0, False), ("mixed", gate, 0.0, False),
                                ("density+mixed", gate, tau, False), ("learned+mixed", gate, tau, True)):
        _lib.check(lib.tlno_gpu_set_mixed_precision(plan.handle, g))
        _lib.check(lib.tlno_gpu_set_density_threshold(plan.handle, t))
        plan.set_learned_screen(params if use_net else None, margin)
        _lib.check(lib.tlno_gpu_build_jk_device(plan.handle, dD.ptr, dJ.ptr, dK.ptr, thr))
        hipmem.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            _lib.check(lib.tlno_gpu_build_jk_device(plan.handle, dD.ptr, dJ.ptr, dK.ptr, thr))
        hipmem.synchronize()
        sec = (time.perf_counter() - t0) / steps
        sq, pq = plan.last_build_stats()
        runs[name] = {"ms": sec * 1e3, "J": dJ.to_host(), "K": dK.to_host(), "shell_quartets": sq,
                      "prim_quartets": pq, "fp32": plan.last_mixed_stats(), "route": plan.last_route_stats()}
    plan.set_learned_screen(None)
    _lib.check(lib.tlno_gpu_set_density_threshold(plan.handle, 0.0))
    _lib.check(lib.tlno_gpu_set_mixed_precision(plan.handle, 0.0))
    base = runs["fp64"]
    out = {"workload": "water40 (40 waters, 1000 Cartesian functions, Schwarz 1e-12)", "fp32_gate": gate,
           "density_threshold": tau, "learned_margin_decades": margin, "fp64_ms": base["ms"],
           "fp64_shell_quartets": base["shell_quartets"]}
    for name in ("mixed", "density+mixed", "learned+mixed"):
        r = runs[name]
        out[name] = {"ms": r["ms"], "speedup_vs_fp64": base["ms"] / r["ms"],
                     "shell_quartets": r["shell_quartets"],
                     "fp32_share_shell_quartets": r["fp32"]["fp32_shell_quartets"] / max(1, r["shell_quartets"]),
                     "max_abs_err_J": float(np.abs(r["J"] - base["J"]).max()),
                     "max_abs_err_K": float(np.abs(r["K"] - base["K"]).max()),
                     "rms_err_J": float(np.sqrt(np.mean((r["J"] - base["J"]) ** 2))),
                     "rms_err_K": float(np.sqrt(np.mean((r["K"] - base["K"]) ** 2)))}
        if r["route"]["candidates"]:
            out[name]["route"] = r["route"]
    out["learned_gain_over_density_only"] = runs["density+mixed"]["ms"] / runs["learned+mixed"]["ms"]
    out["fp64_kernel_choice"] = {k: v["choice"] for k, v in tuning.items()}
    return out


def water100_run(gate=1e-6, tau=SCREEN_TAU):
    """water100 (100 waters at water40's number density, 2500 Cartesian
    functions), Schwarz 1e-12, densities resident in HBM: the Schwarz-screened
    FP64 build judges the network-routed build (learned+mixed, shipped
    weights) and the same build without the network (density+mixed); one
    build each after a warm-up build of the FP64 plan's kernels."""
    from tlno_amd import learned
    lib = _lib.load()
    b, D = configs.water100()
    thr = 1e-12
    t0 = time.perf_counter()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=thr), cartesian=True)
    plan_s = time.perf_counter() - t0
    M = plan.M
    dD = hipmem.DeviceArray.from_host(D)
    dJ = hipmem.DeviceArray((M, M))
    dK = hipmem.DeviceArray((M, M))
    params, margin = learned.load_weights()
    runs = {}
    for name, g, t, use_net in (("fp64", 0.0, 0.0, False), ("density+mixed", gate, tau, False),
                                ("learned+mixed", gate, tau, True)):
        _lib.check(lib.tlno_gpu_set_mixed_precision(plan.handle, g))
        _lib.check(lib.tlno_gpu_set_density_threshold(plan.handle, t))
        plan.set_learned_screen(params if use_net else None, margin)
        hipmem.synchronize()
        t0 = time.perf_counter()
        _lib.check(lib.tlno_gpu_build_jk_device(plan.handle, dD.ptr, dJ.ptr, dK.ptr, thr))
        hipmem.synchronize()
        sec = time.perf_counter() - t0
        sq, pq = plan.last_build_stats()
        runs[name] = {"ms": sec * 1e3, "J": dJ.to_host(), "K": dK.to_host(), "shell_quartets": sq,
                      "prim_quartets": pq, "fp32": plan.last_mixed_stats(), "route": plan.last_route_stats()}
        print(f"bench: water100 {name} {sec * 1e3:.0f} ms", file=sys.stderr, flush=True)
    plan.set_learned_screen(None)
    base = runs["fp64"]
    out = {"workload": "water100 (100 waters, 2500 Cartesian functions, Schwarz 1e-12)", "fp32_gate": gate,
           "density_threshold": tau, "learned_margin_decades": margin, "plan_seconds": plan_s,
           "fp64_ms": base["ms"], "fp64_shell_quartets": base["shell_quartets"],
           "fp64_prim_quartets": base["prim_quartets"]}
    for name in ("density+mixed", "learned+mixed"):
        r = runs[name]
        out[name] = {"ms": r["ms"], "speedup_vs_fp64": base["ms"] / r["ms"],
                     "shell_quartets": r["shell_quartets"], "prim_quartets": r["prim_quartets"],
                     "fp32_share_shell_quartets": r["fp32"]["fp32_shell_quartets"] / max(1, r["shell_quartets"]),
                     "max_abs_err_J": float(np.abs(r["J"] - base["J"]).max()),
                     "max_abs_err_K": float(np.abs(r["K"] - base["K"]).max()),
                     "rms_err_J": float(np.sqrt(np.mean((r["J"] - base["J"]) ** 2))),
                     "rms_err_K": float(np.sqrt(np.mean((r["K"] - base["K"]) ** 2))), "route": r["route"]}
    out["learned_gain_over_density_only"] = runs["density+mixed"]["ms"] / runs["learned+mixed"]["ms"]
    out["learned_quartet_reduction_over_density_only"] = (runs["learned+mixed"]["shell_quartets"] /
                                                          max(1, runs["density+mixed"]["shell_quartets"]))
    return out


def high_l_run(kind="f", steps=5, warmup=2, natoms=30, box_A=12.0):
    """f or g shells: organic_f(30) / organic_g(30) (organic30's generator and
    box with an f, or an f and a g, polarization shell on the heavy atoms;
    Cartesian, Schwarz 1e-12), J+K builds from a density resident in HBM after
    the plan autotuner; the class pairs with an f or g shell run the team
    evaluator.  Per class pair: serialized kernel time (HIP events),
    algorithmic flops (tlno_amd.flops) and rate."""
    lib = _lib.load()
    make, shells = {"f": (configs.organic_f, "3s2p1d1f"), "g": (configs.organic_g, "3s2p1d1f1g")}[kind]
    b, D = make(natoms=natoms, box_A=box_A)
    thr = 1e-12
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=thr), cartesian=True)
    plan.autotune(schwarz=thr, repeats=1)
    M = plan.M
    dD = hipmem.DeviceArray.from_host(D)
    dJ = hipmem.DeviceArray((M, M))
    dK = hipmem.DeviceArray((M, M))

    def one():
        _lib.check(lib.tlno_gpu_build_jk_device(plan.handle, dD.ptr, dJ.ptr, dK.ptr, thr))

    for _ in range(warmup):
        one()
    hipmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    hipmem.synchronize()
    sec = (time.perf_counter() - t0) / steps
    sq_all, pq_all = plan.last_build_stats()
    _lib.check(lib.tlno_gpu_set_timing(plan.handle, 1))
    for _ in range(3):
        one()
    n = len(flops.CLASS_L)
    sq = (C.c_longlong * (n * n))()
    pq = (C.c_longlong * (n * n))()
    ms = (C.c_double * (n * n))()
    nb = C.c_longlong()
    _lib.check(lib.tlno_gpu_class_stats_all(plan.handle, sq, pq, ms, C.byref(nb)))
    _lib.check(lib.tlno_gpu_set_timing(plan.handle, 0))
    classes = []
    for X in range(n):
        for Y in range(X + 1):
            i = n * X + Y
            if sq[i] == 0 or max(X, Y) < 6:
                continue
            fl = flops.class_flops(X, Y, pq[i], sq[i])
            t = ms[i] / nb.value * 1e-3
            classes.append({"class": f"({flops.CLASS_NAME[X]}|{flops.CLASS_NAME[Y]})", "shell_quartets": sq[i],
                            "prim_quartets": pq[i], "ms": round(ms[i] / nb.value, 4), "gflop": round(fl / 1e9, 3),
                            "tflops": round(fl / t / 1e12, 2) if t > 0 else 0.0})
    classes.sort(key=lambda c: -c["ms"])
    hi_sq = sum(c["shell_quartets"] for c in classes)
    return {"workload": f"organic_{kind}{natoms} ({natoms} atoms in a {box_A:g} A cube, heavy {shells} / H 2s1p, "
                        f"{M} Cartesian functions, Schwarz 1e-12)",
            "ms_per_build": sec * 1e3, "builds_per_s": 1.0 / sec, "shell_quartets_per_build": sq_all,
            "prim_quartets_per_build": pq_all, "shell_quartets_per_s": sq_all / sec,
            "high_l_share_of_shell_quartets": hi_sq / max(1, sq_all),
            "high_l_class_serialized_ms": sum(c["ms"] for c in classes), "high_l_classes": classes}


def onee_run(repeats=5):
    """One-electron S, T, V (Cartesian, the reference's default screening:
    group-pair decay 1e-15, nuclear primitive pairs 1e-15 / sum|Z|) on
    organic30 (340 functions, 30 nuclei) and water100 (2500 functions, 300
    nuclei): device time of the S/T/V kernels (HIP events) and the wall time
    of the call with the host copies of the three matrices; the reference
    engine's S/T/V of organic30 on one host thread beside it."""
    lib = _lib.load()
    out = {}
    for name, (b, _) in (("organic30", configs.organic30()), ("water100", configs.water100())):
        plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
        Z = np.ascontiguousarray(b.nuclear_charges, dtype=np.float64)
        R = np.ascontiguousarray(b.nuclear_centers, dtype=np.float64)
        M = plan.M
        S, T, V = (np.empty((M, M)) for _ in range(3))
        st = np.zeros(6)
        D = _lib.D

        def call():
            _lib.check(lib.tlno_gpu_onee_screened(plan.handle, len(Z), Z.ctypes.data_as(D), R.ctypes.data_as(D),
                                                  1e-15, 1e-15, S.ctypes.data_as(D), T.ctypes.data_as(D),
                                                  V.ctypes.data_as(D), st.ctypes.data_as(D)))
        call()
        kms, wall = [], []
        for _ in range(repeats):
            t0 = time.perf_counter()
            call()
            wall.append((time.perf_counter() - t0) * 1e3)
            kms.append(float(st[5]))
        out[name] = {"n_basis": M, "nuclei": len(Z), "shell_groups": int(st[0]), "kept_group_pairs": int(st[2]),
                     "nuclear_prim_pairs_kept": int(st[4]), "kernel_ms": float(np.median(kms)),
                     "call_ms_with_host_copies": float(np.median(wall)),
                     "prim_pair_nucleus_terms_per_s": st[4] * len(Z) / (np.median(kms) * 1e-3)}
    try:
        from oracle import oracle as O
        ref = O.Reference()
        b, _ = configs.organic30()
        t0 = time.perf_counter()
        ref.onee(b, b.nuclear_charges, b.nuclear_centers, spherical=False, screening=True)
        out["organic30"]["reference_cpu_ms_1_thread_incl_plan"] = (time.perf_counter() - t0) * 1e3
    except Exception as e:  # reference not built on this host
        out["organic30"]["reference_cpu_ms_1_thread_incl_plan"] = f"unavailable: {e}"
    return out


def routing_profile_run(builds=3, gate=1e-6, tau=SCREEN_TAU):
    """The learned+mixed water40 build alone (shipped weights, FP32 gate 1e-6,
    density threshold 2e-9, Schwarz 1e-12, no autotuning), `builds` times:
    the workload of the routing-kernel profiles."""
    from tlno_amd import learned
    lib = _lib.load()
    b, D = configs.water40()
    plan = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
    dD = hipmem.DeviceArray.from_host(D)
    dJ = hipmem.DeviceArray((plan.M, plan.M))
    dK = hipmem.DeviceArray((plan.M, plan.M))
    params, margin = learned.load_weights()
    plan.set_learned_screen(params, margin)
    _lib.check(lib.tlno_gpu_set_mixed_precision(plan.handle, gate))
    _lib.check(lib.tlno_gpu_set_density_threshold(plan.handle, tau))
    ms = []
    for _ in range(builds):
        hipmem.synchronize()
        t0 = time.perf_counter()
        _lib.check(lib.tlno_gpu_build_jk_device(plan.handle, dD.ptr, dJ.ptr, dK.ptr, 1e-12))
        hipmem.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    return {"workload": "water40 learned+mixed (routing profile)", "builds": builds, "ms": ms,
            "shell_quartets": plan.last_build_stats()[0], "route": plan.last_route_stats()}


def batched_run(steps=5):
    """batch256: 256 molecules of 3-40 atoms (25-460 Cartesian functions,
    every class up to (dd|dd)), FP64, Schwarz 1e-12: one batched build = J and
    K of all 256 molecules.  Reports default launch shapes, autotuned shapes,
    autotuned + HIP-graph replay, and 256 separate single-molecule builds."""
    from tlno_amd.batch import BatchPlan
    lib = _lib.load()
    mols = configs.batch256()
    t0 = time.perf_counter()
    bp = BatchPlan([b for b, _ in mols])
    plan_s = time.perf_counter() - t0
    dD = hipmem.DeviceArray.from_host(bp.pack([D for _, D in mols]))
    dJ = hipmem.DeviceArray((bp.total,))
    dK = hipmem.DeviceArray((bp.total,))

    def timed(use_graph):
        bp.jk_device(dD, dJ, dK, use_graph=use_graph)
        t = time.perf_counter()
        for _ in range(steps):
            bp.jk_device(dD, dJ, dK, use_graph=use_graph)
        return (time.perf_counter() - t) / steps

    t_default = timed(False)
    ref_J = dJ.to_host()
    tune = bp.autotune(repeats=2)
    t_tuned = timed(False)
    t_graph = timed(True)
    st = bp.last_stats()
    err = float(np.abs(dJ.to_host() - ref_J).max())
    # the same molecules one plan per molecule, device-resident densities
    singles = []
    for b, D in mols:
        p = DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True)
        singles.append((p, hipmem.DeviceArray.from_host(D), hipmem.DeviceArray(D.shape), hipmem.DeviceArray(D.shape)))
    for p, d, j, k in singles:
        _lib.check(lib.tlno_gpu_build_jk_device(p.handle, d.ptr, j.ptr, k.ptr, 1e-12))
    t = time.perf_counter()
    for p, d, j, k in singles:
        _lib.check(lib.tlno_gpu_build_jk_device(p.handle, d.ptr, j.ptr, k.ptr, 1e-12))
    hipmem.synchronize()
    t_single = time.perf_counter() - t
    return {"workload": "batch256: 256 molecules, 3-40 atoms, heavy 3s2p1d / H 2s1p Cartesian, Schwarz 1e-12",
            "molecules": bp.n_mol, "functions_min_max": [int(bp.M.min()), int(bp.M.max())],
            "plan_seconds": plan_s, "shell_quartets_per_build": st["shell_quartets"],
            "prim_quartets_per_build": st["prim_quartets"],
            "ms_per_batched_build": {"default_shapes": t_default * 1e3, "autotuned": t_tuned * 1e3,
                                     "autotuned_graph": t_graph * 1e3, "256_single_molecule_builds": t_single * 1e3},
            "batched_builds_per_s": 1.0 / t_graph, "molecule_builds_per_s": bp.n_mol / t_graph,
            "graph_vs_default_max_abs_diff_J": err, "autotuner": tune}


def host_cpu():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(b, rounds=3):
    """The reference engine (oracle/_ref, built from its own sources) on this
    host: after one warm-up build, `rounds` interleaved rounds of a full
    G = J - K/2 build on one thread (its K and G endpoints refuse more,
    rys/native_quartets.ipp:666-670) and a J-only build on all the host's
    cores (its OpenMP J, rys/native_parallel_j.ipp:12-114); medians reported
    (benchmarks/two_electron_engine/scripts/benchmark_statistics.py:99-131)."""
    try:
        from oracle import oracle as O
        ref = O.Reference()
    except Exception as e:  # reference not built on this host
        return {"value": None, "unit": "Fock builds/s", "cores": 0, "kind": "reference",
                "sample": f"unavailable: {e}"}
    M = ref.plan_M(b)
    D = configs.spherical_density(M, 5, 60)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    cores = max(1, min(cores, os.cpu_count() or 1))
    ref.direct(b, D, "J", schwarz=1e-12, threads=cores)  # warm-up (page-in, thread pool)
    g_s, j_s = [], []
    for _ in range(rounds):
        g_s.append(ref.direct(b, D, "G", schwarz=1e-12, threads=1)[1])
        j_s.append(ref.direct(b, D, "J", schwarz=1e-12, threads=cores)[1])
    g = float(np.median(g_s))
    j = float(np.median(j_s))
    ncpu = os.cpu_count()
    return {"value": 1.0 / g, "unit": "Fock builds/s", "cores": 1, "kind": "reference",
            "sample": (f"full G = J - K/2 builds of the same molecule and basis in the reference's spherical basis "
                       f"(M={M}; its public 2e API has no Cartesian d), Schwarz 1e-12, 1 thread: median of "
                       f"{rounds} interleaved builds after a warm-up, {g:.2f} s (all: "
                       f"{', '.join(f'{x:.2f}' for x in g_s)}); host {host_cpu()}, {ncpu} cpus"),
            "j_only_all_cores": {"value": 1.0 / j, "unit": "J builds/s", "cores": cores,
                                 "median_s": j, "all_s": j_s}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="organic30", choices=["organic30", "water40"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-water100", action="store_true")
    ap.add_argument("--no-onee", action="store_true")
    ap.add_argument("--no-fshell", action="store_true")
    ap.add_argument("--no-gshell", action="store_true")
    ap.add_argument("--no-queue-probe", action="store_true",
                    help="skip the organic30 rerun with HIP's default 4 hardware queues")
    ap.add_argument("--routing-profile", type=int, default=0, metavar="BUILDS",
                    help="only run BUILDS learned+mixed water40 builds (the routing-kernel profile workload)")
    ap.add_argument("--onee-profile", action="store_true",
                    help="only run the S/T/V leg (organic30 and water100) and print it as one JSON line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy a buffer of this many MiB device to device (a copy kernel of known traffic "
                         "for calibrating hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the J/K build shards by "
              "quartet tasks with one all-reduce of J and K, not built here)", file=sys.stderr)
        sys.exit(2)
    thr = 1e-12
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src, dst = hipmem.DeviceArray((n,)), hipmem.DeviceArray((n,))
        dst.copy_from(src)
        hipmem.synchronize()
        src.free()
        dst.free()
    if args.routing_profile > 0:
        print(json.dumps(routing_profile_run(args.routing_profile)), flush=True)
        return
    if args.onee_profile:
        print(json.dumps(onee_run()), flush=True)
        return
    b, D = configs.organic30() if args.workload == "organic30" else configs.water40()
    print(f"bench: {args.workload} J/K builds ({args.warmup} warm-up, {args.steps} timed)", file=sys.stderr,
          flush=True)
    plan, wall, classes, J, K = gpu_run(b, D, args.steps, args.warmup, thr)
    ms_per = wall / args.steps * 1e3
    sq_tot = sum(c["shell_quartets"] for c in classes)
    pq_tot = sum(c["prim_quartets"] for c in classes)
    ints_tot = sum(c["integrals"] for c in classes)
    dom = max(classes, key=lambda c: c["ms"])
    dom_kernel = (f"k_eri_team {dom['class']}" if plan.tuning.get(dom["class"], {}).get("choice") == "team"
                  else f"k_eri_jk {dom['class']} (all chunk groups)")
    total_gflop = sum(c["gflop"] for c in classes)
    result = {
        "metric": "J/K Fock builds/s",
        "value": args.steps / wall,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, tlno_amd.configs)",
        "config": {"workload": f"{args.workload}: direct J+K build, Cartesian d, Schwarz 1e-12",
                   "n_basis": plan.M, "shell_pairs": plan.profile.pair_groups,
                   "primitive_pairs": plan.profile.primitive_pairs, "parallelism": "single GPU",
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
        "eri_throughput": {"shell_quartets_per_s": sq_tot / (wall / args.steps),
                           "primitive_quartets_per_s": pq_tot / (wall / args.steps),
                           "cartesian_integrals_per_s": ints_tot / (wall / args.steps),
                           "shell_quartets_per_build": sq_tot, "primitive_quartets_per_build": pq_tot},
        "roofline": {"bound": "fp64_valu", "pipe": "fp64 VALU (MI355X fp64 vector peak = fp64 matrix peak)",
                     "kernel": dom_kernel, "achieved": dom["tflops"],
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": dom["tflops"] / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic(dom_kernel), "algorithmic_gflop_per_launch": dom["gflop"],
                     "avg_launch_ms": dom["ms"],
                     "whole_build_tflops": total_gflop / (ms_per * 1e-3) / 1e3},
        "classes": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in c.items()} for c in classes],
        "autotuner": {"seconds": plan.tuning_seconds, "classes": plan.tuning},
    }
    if not args.no_mixed:
        print("bench: water40 mixed-precision and screened builds", file=sys.stderr, flush=True)
        result["mixed_precision"] = mixed_run()
    if not args.no_water100:
        print("bench: water100 learned-routed builds", file=sys.stderr, flush=True)
        result["learned_water100"] = water100_run()
    if not args.no_onee:
        print("bench: one-electron S/T/V", file=sys.stderr, flush=True)
        result["onee"] = onee_run()
    if not args.no_fshell:
        print("bench: f-shell builds (organic_f30)", file=sys.stderr, flush=True)
        result["fshell"] = high_l_run("f")
    if not args.no_gshell:
        print("bench: g-shell builds (organic_g30)", file=sys.stderr, flush=True)
        result["gshell"] = high_l_run("g")
    if not args.no_batch:
        print("bench: 256-molecule batch builds", file=sys.stderr, flush=True)
        result["batched"] = batched_run()
    if not args.no_queue_probe:
        print("bench: organic30 with 4 hardware queues (HIP's default)", file=sys.stderr, flush=True)
        result["default_hw_queues"] = queue_probe(args.steps, args.warmup)
    if not args.no_cpu:
        print("bench: reference CPU build", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(b)
        if result["cpu_baseline"]["value"]:
            result["gpu_over_cpu_baseline"] = result["value"] / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


HW_QUEUES = 8


def queue_probe(steps, warmup, queues=4):
    """The headline build in a child process with `queues` hardware queues
    (HIP's default is 4; the headline runs with 8): builds/s and ms."""
    import subprocess
    env = dict(os.environ, GPU_MAX_HW_QUEUES=str(queues), TLNO_BENCH_QUEUE_PROBE="1")
    cmd = [sys.executable, os.path.abspath(__file__), "--no-cpu", "--no-mixed", "--no-batch", "--no-water100",
           "--no-onee", "--no-queue-probe", "--steps", str(steps), "--warmup", str(warmup)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"hw_queues": queues, "error": r.stderr[-400:]}
    line = json.loads(r.stdout.strip().splitlines()[-1])
    return {"hw_queues": queues, "value": line["value"], "ms_per_step": line["ms_per_step"]}


def relaunch_with_hw_queues():
    """The plan runs the class kernels of a build on up to 8 concurrent
    streams, which HIP spreads over GPU_MAX_HW_QUEUES hardware queues (4 by
    default), read when the HIP runtime loads.  The benchmark runs with 8
    queues: if the environment says otherwise it reruns itself as a child
    process with GPU_MAX_HW_QUEUES=8 (nothing has touched the GPU yet) and
    exits with the child's status."""
    if os.environ.get("GPU_MAX_HW_QUEUES") == str(HW_QUEUES) or os.environ.get("TLNO_BENCH_QUEUE_PROBE") == "1":
        return
    import subprocess
    env = dict(os.environ, GPU_MAX_HW_QUEUES=str(HW_QUEUES))
    sys.exit(subprocess.run([sys.executable, os.path.abspath(__file__)] + sys.argv[1:], env=env).returncode)


if __name__ == "__main__":
    relaunch_with_hw_queues()
    main()
