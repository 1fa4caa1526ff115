"""Trains the learned shell-quartet screen on the GPU and writes its weights.

    python -m tlno_amd.train_screen [--out PATH] [--molecules 500] [--per-molecule 4000]

Training set: per_molecule Schwarz-surviving quartets from each of
n_molecules generated molecules (3-40 atoms, heavy 3s2p1d / H 2s1p, three
primitives per shell, one basis set per molecule), labels from the engine.
Symmetry zeros are left out, 10% of the molecules are held out, and the
safety margin is the 0.9999 quantile of the held-out under-prediction plus
0.25 decades.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import learned


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=learned.DEFAULT_WEIGHTS)
    ap.add_argument("--molecules", type=int, default=500)
    ap.add_argument("--per-molecule", type=int, default=4000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--w-under", type=float, default=32.0)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args(argv)
    t0 = time.time()
    ts = learned.make_training_set(args.molecules, args.per_molecule,
                                   progress=lambda i: print(f"labelled {i} molecules ({time.time() - t0:.0f} s)",
                                                            flush=True))
    t_data = time.time() - t0
    hold = ts.molecule >= int(0.9 * args.molecules)
    use = ts.trainable()
    Xtr, ttr = ts.X[~hold & use], ts.t[~hold & use].astype(np.float32)
    Xva, tva = ts.X[hold & use], ts.t[hold & use]
    net = learned.ScreenNetwork(seed=args.seed)
    net.set_normalization(Xtr)
    t1 = time.time()
    hist = net.train(Xtr, ttr, epochs=args.epochs, batch=args.batch, lr=args.lr, w_under=args.w_under,
                     seed=args.seed)
    t_train = time.time() - t1
    for e, l in enumerate(hist):
        print(f"epoch {e:3d} loss {l:.5f}", flush=True)
    yva = net.predict(Xva, matrix_cores=True).astype(np.float64)
    res = tva - yva
    margin = learned.choose_margin(res)
    stats = {
        "samples_train": int(len(Xtr)), "samples_heldout": int(len(Xva)),
        "symmetry_zeros_left_out": int((~ts.trainable()).sum()), "w_under": args.w_under, "epochs": args.epochs,
        "data_s": round(t_data, 1), "train_s": round(t_train, 2),
        "heldout_rmse": float(np.sqrt(np.mean(res ** 2))),
        "heldout_under_q9999": float(np.quantile(res, 0.9999)),
        "heldout_max_under": float(res.max()),
        "margin": margin,
        "label_mean": float(np.mean(tva)),
        # share of held-out quartets whose bound tightens by at least 10x
        "tighter_10x": float(np.mean(np.minimum(0.0, yva + margin) <= -1.0)),
    }
    print(json.dumps(stats), flush=True)
    learned.save_weights(args.out, net.params, margin, loss_history=hist,
                         stats=json.dumps(stats))
    return 0


if __name__ == "__main__":
    sys.exit(main())
