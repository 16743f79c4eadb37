"""Coefficient tables: parsing, Brent equations, checksums, CSE expansion.

Pins the table reader (codegen/gen.py:101-170 grammar and grouping) with
properties the mathematics fixes: an exact algorithm satisfies the Brent
equations exactly (codegen/verify.py:110-132), an APA one up to O(x); the
nonzero count is the number in the table's file name; an eliminated table
(codegen/subexpr_elim.py:186-213) expands back to its original.
"""
import os
import re
from fractions import Fraction

import pytest

from fastmm import tables as T

EXACT = [n for n in T.algorithm_names() if not T.get_table(n).is_apa]
APA = [n for n in T.algorithm_names() if T.get_table(n).is_apa]


@pytest.mark.parametrize("name", EXACT)
def test_brent_exact(name):
    t = T.get_table(name)
    assert T.brent_residual(t) == 0


@pytest.mark.parametrize("name", APA)
def test_brent_apa_is_order_x(name):
    t = T.get_table(name)
    r1 = T.brent_residual(t, Fraction(1, 10**6))
    r2 = T.brent_residual(t, Fraction(1, 10**8))
    assert 0 < r1 < Fraction(1, 10**4)
    assert r2 < r1 / 50  # residual shrinks (at least) linearly with x


@pytest.mark.parametrize("name", [n for n in T.algorithm_names()
                                  if n in T.derived_algorithms()
                                  or not T._REGISTRY[n][0].startswith("eliminated/")])
def test_nnz_matches_file_name(name):
    t = T.get_table(name)
    nums = re.findall(r"\d+", os.path.basename(t.path))
    if not nums:
        assert t.nnz() == 36  # strassen: 12 nonzeros in each of U, V, W
        return
    assert str(t.nnz()) == nums[-1]
    if name in T.derived_algorithms():  # the derived name carries the same count
        assert name.endswith("_" + nums[-1])


PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


@pytest.mark.parametrize("name", ["strassen", "grey322_11_50", "grey423_20_144", "grey424_26_257",
                                  "smirnov333_23_139", "hk332_15_94", "grey522_18_99"])
def test_all_six_permutations_are_exact(name):
    """codegen/convert.py:61-99 derives every <.,.,.> order of a table; each
    must satisfy the Brent equations exactly, with dims permuted and the
    coefficient multiset unchanged."""
    t = T.get_table(name)
    d = (t.m, t.k, t.n)
    for perm in PERMS:
        p = T.permute_table(t, perm)
        assert (p.m, p.k, p.n) == (d[perm[0]], d[perm[1]], d[perm[2]])
        assert p.R == t.R and p.nnz() == t.nnz()
        assert T.brent_residual(p) == 0, perm


def test_permutation_group_laws():
    """Transposition is an involution and the cyclic shift has order 3."""
    t = T.get_table("grey423_20_144")
    same = lambda a, b: ((a.m, a.k, a.n), a.U, a.V, a.W) == ((b.m, b.k, b.n), b.U, b.V, b.W)
    assert same(T.permute_table(T.permute_table(t, (2, 1, 0)), (2, 1, 0)), t)
    c = T.permute_table(t, (2, 0, 1))
    c = T.permute_table(c, (2, 0, 1))
    c = T.permute_table(c, (2, 0, 1))
    assert same(c, t)


def test_brent_check_catches_a_wrong_permutation():
    """A plausible slip (transposing without moving W's block grid) breaks
    the Brent equations, so the exactness check above has teeth."""
    t = T.get_table("grey423_20_144")
    good = T.permute_table(t, (2, 1, 0))
    bad = T.Table(name="bad", m=good.m, k=good.k, n=good.n, U=good.U, V=good.V,
                  W=[list(r) for r in t.W])   # W rows left in <4,2,3> order
    assert T.brent_residual(bad) != 0
    with pytest.raises(ValueError):
        T.permute_table(T.get_table("grey424_26_206"), (1, 0, 2))  # CSE groups


@pytest.mark.parametrize("elim,orig", [("grey424_26_206", "grey424_26_257"),
                                       ("grey323_15_89", "grey323_15_103"),
                                       ("grey333_23_125", None), ("hk323_15_84", "hk323_15_94"),
                                       ("smirnov333_23_128", None)])
def test_eliminated_expands_to_original(elim, orig):
    e = T.get_table(elim)
    Ue, Ve, We = e.expanded()
    # the expansion is itself an exact algorithm
    assert T.brent_residual(e) == 0
    if orig is None:
        return
    o = T.get_table(orig)
    Uo, Vo, Wo = o.expanded()
    assert (Ue, Ve, We) == (Uo, Vo, Wo)


def test_strassen_known_table():
    t = T.get_table("strassen")
    assert (t.m, t.k, t.n, t.R) == (2, 2, 2, 7)
    # M1 = (A11 + A22)(B11 + B22), the textbook first product
    assert [p for p, _ in t.s_terms(0)] == [0, 3]
    assert [p for p, _ in t.t_terms(0)] == [0, 3]


def test_parse_coeff_grammar():
    x = Fraction(1, 100)
    assert T.parse_coeff("-.5").value == Fraction(-1, 2)
    assert T.parse_coeff("1/2").value == Fraction(1, 2)
    assert T.parse_coeff("x").eval_exact(x) == x
    assert T.parse_coeff("-x").eval_exact(x) == -x
    assert T.parse_coeff("xi").eval_exact(x) == 1 / x
    assert T.parse_coeff("x2").eval_exact(x) == x * x
    assert T.parse_coeff("-x2i").eval_exact(x) == -1 / (x * x)
    assert T.parse_coeff("0.5x").eval_exact(x) == x / 2


def test_blank_line_keeps_empty_group_slot():
    # eliminated/grey322-11-49 writes an empty A-side substitution group as a
    # blank line; its B-side group must still land in slot 4
    t = T.load_table(os.path.join(T.TABLE_DIR, "eliminated", "grey322-11-49"))
    assert len(t.SX) == 0 and len(t.TX) == 1
    assert T.brent_residual(t) == 0


def test_enum_registry_quirks():
    # FAST323_15_89 dispatches to grey323_15_103 (algorithms/all_algorithms.hpp:159-160)
    assert T.ENUM_TO_NAMESPACE["FAST323_15_89"] == "grey323_15_103"
    assert T.ENUM_TO_NAMESPACE["MKL"] == "classical222_8_24"
    assert T.default_x("BINI322_10_52_APPROX", 2) == 1e-4
    assert T.default_x("SCHONHAGE333_21_117_APPROX", 2) == 1e-2
