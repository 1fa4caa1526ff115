"""GPU: the reference's one-electron boundary re-exported over the GPU
engine, against the reference's own C ABI outputs on its own test molecules
(tests/golden/onee_capi.npz: h2o.inp with STO-3G, gly.inp with def2-SVP):

  * tlno_onee_* C ABI (include/tlno_onee.h): AO counts, group-pair and
    nuclear primitive-pair statistics (exact), Cartesian and spherical S/T/V,
    screening on (the reference's default) and off, cached and full plans;
  * the block-level C++ API (onee_integrals.hpp): every group pair's
    contracted blocks, accumulated into caller blocks and scattered, against
    the matrix API and the reference.

Tolerance: |d| <= 1e-10 for S and V; T within 1e-10 as well, except where
the reference's one-sided kinetic formula loses digits (documented in the
oracle tests): there the bound is relative to the largest kinetic entry.
"""
import json
import subprocess

import numpy as np
import pytest

import onee_capi_io as io
from tlno_amd.basis import Basis

pytestmark = pytest.mark.gpu


def _close(a, b, name, tol=1e-10):
    d = float(np.abs(a - b).max())
    assert d <= tol, (name, d)
    return d


@pytest.mark.parametrize("mol,basis", io.CASES)
@pytest.mark.parametrize("screen", ["on", "off"])
def test_onee_capi_matches_reference(mol, basis, screen, tmp_path):
    g = io.gold()
    L = io.lib()
    inp, bas = io.write_inputs(str(tmp_path), mol, basis)
    L.tlno_onee_set_screening(1 if screen == "on" else 0)
    try:
        h = L.tlno_onee_create_from_inp(inp.encode(), bas.encode())
        assert h
        n, ns = L.tlno_onee_nbf(h), L.tlno_onee_nbf_sph(h)
        assert [n, ns] == g[f"{mol}_{screen}_nbf"].tolist()
        got = io.evaluate(h, n, ns, cached=True)
        full = io.evaluate(h, n, ns, cached=False)
        L.tlno_onee_destroy(h)
    finally:
        L.tlno_onee_set_screening(1)
    tag = f"{mol}_{screen}"
    assert got["pair_stats"].tolist() == g[f"{tag}_pair_stats"].tolist()
    assert got["nuclear_pair_stats"].tolist() == g[f"{tag}_nuclear_pair_stats"].tolist()
    diffs = {}
    for k in ("S", "V", "Ssph", "Vsph"):
        diffs[k] = _close(got[k], g[f"{tag}_{k}"], k)
    for k in ("T", "Tsph"):
        diffs[k] = float(np.abs(got[k] - g[f"{tag}_{k}"]).max())
        assert diffs[k] <= 1e-10 * max(1.0, float(np.abs(g[f"{tag}_{k}"]).max())), (k, diffs[k])
    for k in ("S", "T", "V", "Ssph", "Tsph", "Vsph"):
        _close(full[k], got[k], k + " full vs cached", 1e-13)
    print(f"{mol}/{basis} screening {screen}: max|d| " + " ".join(f"{k} {v:.1e}" for k, v in diffs.items()))


def _component_basis(mol, basis):
    g = io.gold()
    lines = str(g[f"{mol}_inp"]).splitlines()
    start = next(i for i, ln in enumerate(lines) if ln.strip().startswith("*"))
    atoms = []
    for ln in lines[start + 1:]:
        t = ln.split()
        if not t or t[0].startswith("*"):
            break
        atoms.append((t[0], np.array([float(x) for x in t[1:4]])))
    el = {k[len(basis) + 1:-5]: json.loads(str(g[k])) for k in g.files
          if k.startswith(basis + "_") and k.endswith("_json")}
    shells, Z, R = [], [], []
    for sym, xyz in atoms:
        e = el[sym]
        Z.append(e["charge"])
        R.append(xyz)
        for sh in e["shells"]:
            prim = np.array(sh[1:], dtype=float)
            shells.append((xyz, sh[0], prim[:, 0], prim[:, 1]))
    return Basis.from_shells(shells, Z, R)


def test_block_api_matches_matrix_api_and_reference(tmp_path):
    b = _component_basis("gly", "def2-svp")
    g = io.gold()
    path = tmp_path / "in.txt"
    with open(path, "w") as f:
        f.write(f"{b.n}\n")
        off = 0
        for i in range(b.n):
            k = int(b.nprim[i])
            prims = " ".join(f"{float(b.exps[off + j])!r} {float(b.coefs[off + j])!r}" for j in range(k))
            c, a = [float(x) for x in b.centers[i]], [int(x) for x in b.ang[i]]
            f.write(f"{c[0]!r} {c[1]!r} {c[2]!r} {a[0]} {a[1]} {a[2]} {k} {prims}\n")
            off += k
        f.write(f"{len(b.nuclear_charges)}\n")
        for z, r in zip(b.nuclear_charges, b.nuclear_centers):
            f.write(f"{int(z)} {float(r[0])!r} {float(r[1])!r} {float(r[2])!r}\n")
    out = tmp_path / "out.txt"
    drv = io.os.path.join(io.ROOT, "tests", "cpp", "onee_block_driver")
    r = subprocess.run([drv, str(path), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = {}
    for line in out.read_text().splitlines():
        name, rest = line.split(" ", 1)
        res[name] = np.array([float(x) for x in rest.split()[1:]])
    n = b.n
    assert res["plan"].tolist()[:3] == [45, 1035, 1035]
    assert int(res["plan"][3]) == 95
    for k in "STV":
        blk = res[f"{k}_blocks"].reshape(n, n)
        _close(blk, res[k].reshape(n, n), f"{k} blocks vs matrix API", 1e-13)
        ref = g[f"gly_off_{k}"]
        tol = 1e-10 * (max(1.0, float(np.abs(ref).max())) if k == "T" else 1.0)
        _close(blk, ref, f"{k} blocks vs reference", tol)
