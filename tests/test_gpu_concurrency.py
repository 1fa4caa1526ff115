"""GPU: builds of different plans and a molecule batch issued at the same time
from several host threads.  The class launches of every build share the
device's side streams (fork/join under a per-device lock); each thread's
results must equal the same builds done one after another (1e-12: only the
order of the floating-point atomics differs)."""
import threading

import numpy as np
import pytest

from tlno_amd import configs
from tlno_amd.batch import BatchPlan
from tlno_amd.fock import DirectPlan, ScreeningOptions

pytestmark = pytest.mark.gpu


def test_threads_building_different_plans_and_a_batch():
    cases = [configs.water_cluster(5), configs.batch256(seed=11, nmol=1, atom_range=(9, 9))[0]]
    plans = [DirectPlan(b, ScreeningOptions(integral_threshold=1e-12), cartesian=True) for b, _ in cases]
    dens = [configs.spherical_density(p.M, 30 + i, 12) for i, p in enumerate(plans)]
    mols = configs.batch256(seed=12, nmol=5, atom_range=(3, 8))
    bp = BatchPlan([b for b, _ in mols])
    bdens = [D for _, D in mols]

    ref_plans = [p.jk(D) for p, D in zip(plans, dens)]
    ref_batch = bp.jk(bdens)

    out, errors = {}, []

    def plan_worker(i):
        try:
            out[i] = [plans[i].jk(dens[i]) for _ in range(4)]
        except Exception as e:  # reported below
            errors.append(repr(e))

    def batch_worker():
        try:
            out["batch"] = [bp.jk(bdens) for _ in range(4)]
        except Exception as e:
            errors.append(repr(e))

    threads = [threading.Thread(target=plan_worker, args=(i,)) for i in range(len(plans))]
    threads.append(threading.Thread(target=batch_worker))
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors

    worst = 0.0
    for i, (J0, K0) in enumerate(ref_plans):
        for J, K in out[i]:
            worst = max(worst, np.abs(J - J0).max(), np.abs(K - K0).max())
    for Js, Ks in out["batch"]:
        for J, K, J0, K0 in zip(Js, Ks, *ref_batch):
            worst = max(worst, np.abs(J - J0).max(), np.abs(K - K0).max())
    print(f"threads: max |d| against sequential builds {worst:.2e}")
    assert worst < 1e-12
