"""CPU: the reference-compatible one-electron C ABI (include/tlno_onee.h) is
exported by the facade library, reads the reference's molecule format and
JSON basis files into the same basis (the AO count and the exported basis
parameters equal the reference's on its own test molecules), and returns
nullptr / -1 where the reference does.  No GPU work."""
import json
import re

import pytest

import onee_capi_io as io

HEADER = io.os.path.join(io.ROOT, "include", "tlno_onee.h")


def test_facade_exports_every_onee_symbol():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(tlno_onee_\w+)\s*\(", text)))
    assert len(names) == 22
    L = io.lib()
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("mol,basis", io.CASES)
def test_inp_and_json_basis_match_reference_counts(mol, basis, tmp_path):
    g = io.gold()
    inp, bas = io.write_inputs(str(tmp_path), mol, basis)
    L = io.lib()
    h = L.tlno_onee_create_from_inp(inp.encode(), bas.encode())
    assert h
    try:
        assert L.tlno_onee_nbf(h) == int(g[f"{mol}_on_nbf"][0])
    finally:
        L.tlno_onee_destroy(h)
    # the exported parameters round-trip to the reference's own text
    for k in g.files:
        if k.startswith(basis + "_") and k.endswith("_json"):
            el = k[len(basis) + 1:-5]
            assert json.loads(L.tlno_onee_basis_json(bas.encode(), el.encode()).decode()) == json.loads(str(g[k]))
            assert L.tlno_onee_basis_json(bas.encode(), el.encode()).decode() == str(g[k])


def test_reference_failure_behaviour(tmp_path):
    inp, bas = io.write_inputs(str(tmp_path), "h2o", "sto-3g")
    L = io.lib()
    assert not L.tlno_onee_create_from_inp(inp.encode(), b"sto-3g")  # named tables are not built
    assert not L.tlno_onee_create_from_inp(b"/nonexistent.inp", bas.encode())
    bad = tmp_path / "bad.inp"
    bad.write_text("* xyz 0 1\nO 0 0 0\n")  # unterminated block
    assert not L.tlno_onee_create_from_inp(str(bad).encode(), bas.encode())
    assert L.tlno_onee_nbf(None) == -1 and L.tlno_onee_nbf_sph(None) == -1
    L.tlno_onee_noop(None)
    L.tlno_onee_destroy(None)
    assert not L.tlno_onee_basis_json(b"def2-svp", b"O")
    L.tlno_onee_set_screening(0)
    assert L.tlno_onee_screening_enabled() == 0
    assert L.tlno_onee_group_pair_decay_eps() == 0.0 and L.tlno_onee_nuclear_screen_eps() == 0.0
    L.tlno_onee_set_screening(1)
    assert L.tlno_onee_screening_enabled() == 1
    assert L.tlno_onee_group_pair_decay_eps() == 1e-15 and L.tlno_onee_nuclear_screen_eps() == 1e-15
