"""Input writer / output reader for tests/cpp/facade_driver (the C++ facade
exercised as a reference caller would use it)."""
import os
import subprocess

import numpy as np

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_driver")


def write_input(path, b, D0, D1):
    with open(path, "w") as f:
        n = len(b.nprim)
        f.write(f"{n}\n")
        off = 0
        for i in range(n):
            c, a, k = b.centers[i], b.ang[i], int(b.nprim[i])
            prims = " ".join(f"{float(b.exps[off + j])!r} {float(b.coefs[off + j])!r}" for j in range(k))
            f.write(f"{float(c[0])!r} {float(c[1])!r} {float(c[2])!r} {int(a[0])} {int(a[1])} {int(a[2])} {k} {prims}\n")
            off += k
        f.write(f"{len(b.nuclear_charges)}\n")
        for z, r in zip(b.nuclear_charges, b.nuclear_centers):
            f.write(f"{int(z)} {float(r[0])!r} {float(r[1])!r} {float(r[2])!r}\n")
        f.write(f"{D0.shape[0]}\n")
        for D in (D0, D1):
            f.write(" ".join(repr(float(x)) for x in D.ravel()) + "\n")


def run(b, D0, D1, workdir):
    inp, out = os.path.join(workdir, "in.txt"), os.path.join(workdir, "out.txt")
    write_input(inp, b, D0, D1)
    r = subprocess.run([DRIVER, inp, out], capture_output=True, text=True, timeout=300)
    res = {}
    if r.returncode == 0:
        with open(out) as f:
            for line in f:
                name, rest = line.split(" ", 1)
                if name in ("error", "inc_error"):
                    res.setdefault(name, []).append(rest.strip())
                    continue
                vals = rest.split()
                res[name] = np.array([float(x) for x in vals[1:]])
    return r, res
