"""The table reader and the HIP generator's operand maps against the
structure of the reference's committed generated headers.

header_structure.json was extracted from algorithms/*.hpp by
tests/golden/extract_header_structure.py (S_Add/T_Add/M_Add/SX/TX/MX bodies
and call sites, the operands of each FastMatmulRecursive call and the
UpdateMultiplier coefficients).  Every product's S and T operand, every
C block's W row and every CSE temporary read from our tables must match it
term for term, in the same order (the order fixes the rounding).
"""
import json
import os
import subprocess
import sys
from fractions import Fraction

import pytest

from fastmm import tables as T

HERE = os.path.dirname(os.path.abspath(__file__))
STRUCT = json.load(open(os.path.join(HERE, "golden", "header_structure.json")))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "fast-matmul_amd", "gen"))
import hipgen  # noqa: E402


def _names(t):
    def a(p):
        if p < t.m * t.k:
            return "A%d%d" % (p // t.k + 1, p % t.k + 1)
        return "A_X%d" % (p - t.m * t.k + 1)

    def b(p):
        if p < t.k * t.n:
            return "B%d%d" % (p // t.n + 1, p % t.n + 1)
        return "B_X%d" % (p - t.k * t.n + 1)

    def mm(q):
        return "M%d" % (q + 1) if q < t.R else "M_X%d" % (q - t.R + 1)

    return a, b, mm


def _terms(lst, namef):
    return [[namef(i), Fraction(c.value)] for i, c in lst]


def _fix(lst):
    return [[n, Fraction(c)] for n, c in lst]


@pytest.mark.parametrize("name", sorted(STRUCT))
def test_structure_matches_reference_headers(name):
    t = T.get_table(name)
    ref = STRUCT[name]
    an, bn, mn = _names(t)
    assert len(ref["products"]) == t.R
    for r, pr in enumerate(ref["products"]):
        assert pr["M"] == "M%d" % (r + 1)
        s, tt = t.s_terms(r), t.t_terms(r)
        mult = Fraction(1)
        if len(s) > 1:
            assert _fix(pr["S"]) == _terms(s, an), "S%d" % (r + 1)
        else:
            assert pr["S"][0][0] == an(s[0][0])
            mult *= s[0][1].value
        if len(tt) > 1:
            assert _fix(pr["T"]) == _terms(tt, bn), "T%d" % (r + 1)
        else:
            assert pr["T"][0][0] == bn(tt[0][0])
            mult *= tt[0][1].value
        assert Fraction(pr["mult"]) == mult
    for c in range(t.m * t.n):
        key = "C%d%d" % (c // t.n + 1, c % t.n + 1)
        assert _fix(ref["C"][key]) == _terms(t.w_terms(c), mn), key
    for i in range(len(t.SX)):
        assert _fix(ref["SX"]["A_X%d" % (i + 1)]) == _terms(t.sx_terms(i), an)
    for i in range(len(t.TX)):
        assert _fix(ref["TX"]["B_X%d" % (i + 1)]) == _terms(t.tx_terms(i), bn)
    for i in range(len(t.MX)):
        if t.mx_terms(i):
            assert _fix(ref["MX"]["M_X%d" % (i + 1)]) == _terms(t.mx_terms(i), mn)


@pytest.mark.parametrize("name", ["strassen", "grey424_26_206", "smirnov333_23_139",
                                  "grey423_20_144", "hk323_15_84", "bini322_10_52_approx"])
def test_generator_operand_map(name):
    """Each product's operand in the generated map is the materialised
    output, X temporary or block that the table names."""
    t = T.get_table(name)
    for side, cols, nblk in (("A", [t.s_terms(r) for r in range(t.R)], t.m * t.k),
                             ("B", [t.t_terms(r) for r in range(t.R)], t.k * t.n)):
        info = hipgen.SideInfo(t, side)
        for r, terms in enumerate(cols):
            kind, idx, coef = info.src[r]
            if len(terms) > 1:
                assert kind == "buf" and info.outputs[idx] == ("col", r)
            elif terms[0][0] < nblk:
                assert kind == "blk" and idx == terms[0][0]
            else:
                assert kind == "buf" and info.outputs[idx] == ("x", terms[0][0] - nblk)


def test_generated_sources_are_current(tmp_path):
    """The committed csrc/gen files are what the generator writes now."""
    out = tmp_path / "gen"
    subprocess.check_call([sys.executable, os.path.join(ROOT, "fast-matmul_amd", "gen", "hipgen.py"),
                           str(out)], stdout=subprocess.DEVNULL)
    gdir = os.path.join(ROOT, "fast-matmul_amd", "csrc", "gen")
    for f in os.listdir(out):
        assert open(os.path.join(out, f)).read() == open(os.path.join(gdir, f)).read(), f
