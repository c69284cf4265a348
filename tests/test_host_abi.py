"""CPU checks of the product library: it loads, exports every symbol the C
header declares, and its host-side coefficient generator (exact rational
Lagrange weights) reproduces the oracle's restatement of the reference
(tables + mapping) bit for bit.  No device compute happens here."""
import re
from pathlib import Path

import numpy as np
import pytest

import impact_hip as ih
import oracle_lib as O

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    text = (ROOT / "include" / "impact_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ih_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = ih.load()
    names = header_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version():
    assert ih.load().ih_abi_version() == ih.IH_ABI_VERSION


COEFF_NAMES = []
for a, b, cc, c2 in (("1", "u", "u1", "u11"), ("2", "v", "v2", "v22"), ("3", "w", "w3", "w33")):
    COEFF_NAMES += [f"cGp{a}", f"cD{cc}", f"cIp{b}", f"cI{b}p", f"cNp{a}D", f"cNp{a}U", f"cN{cc}D", f"cN{cc}U",
                    f"cp{a}{a}", "c" + c2, f"cdg{a}"]

GRIDS = [
    ((17, 13, 11), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5)),
    ((33, 17, 9), (-1, -1, 1, 1, -1, -1), (1.0, 1.0, 1.0), (0.0, 1.2, 0.0)),
    ((65, 65, 65), (1, 1, 1, 1, 1, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)),
    ((257, 129, 129), (1, 1, 1, 1, 1, 1), (2.0, 1.0, 1.0), (0.0, 1.5, 1.5)),
]


@pytest.mark.parametrize("N,bc,L,st", GRIDS)
def test_host_coefficients_match_oracle_bitwise(N, bc, L, st):
    oc = O.OracleCase(N, bc, L, st)
    for name in COEFF_NAMES:
        a = oc.coeff(name)
        b = ih.grid_coeff(N, bc, L, st, name)
        assert a.shape == b.shape and np.array_equal(a, b), name
    for d in range(3):
        for which in (0, 1):
            assert np.array_equal(oc.coords(d, which), ih.grid_coords(N, bc, L, st, d, which))


def test_host_rejects_unsupported_grids():
    with pytest.raises(ih.IHError):
        ih.grid_coeff((17, 9, 9), (0, 0, 1, 1, 1, 1), (1, 1, 1), (0, 0, 0), "cGp1")  # neighbour-block BC
    with pytest.raises(ih.IHError):
        ih.grid_coeff((17, 9, 9), (-1, 1, 1, 1, 1, 1), (1, 1, 1), (0, 0, 0), "cGp1")  # one-sided periodic
    with pytest.raises(ih.IHError):
        ih.grid_coeff((4, 9, 9), (1, 1, 1, 1, 1, 1), (1, 1, 1), (0, 0, 0), "cGp1")  # too small
    with pytest.raises(ih.IHError):
        ih.grid_coeff((17, 9, 9), (1, 1, 1, 1, 1, 1), (1, 1, 1), (0, 0, 0), "cFoo")
    assert "unknown coefficient" in ih.last_error()
