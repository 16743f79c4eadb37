"""Coefficient tables <m,k,n;R> (U, V, W) and the algorithm registry.

A table file is the format read by the reference generator
(codegen/gen.py:26-50 describes it, codegen/gen.py:150-170 reads it):

* group 0: U, one row per A block in row-major block order, one column per
  product r (the S_r = sum_p U[p, r] * A_p side);
* group 1: V, the same for B blocks (T_r);
* group 2: W, one row per C block (row-major), one column per product;
* optional groups 3, 4, 5 (tables written by codegen/subexpr_elim.py:186-213):
  definitions of the common-subexpression temporaries A_X, B_X and M_X.
  When present, U and V carry extra rows (one per A_X / B_X) after the block
  rows and W carries extra columns (one per M_X) after the R product columns.

Lines starting with '#' separate groups; a group is only closed when it holds
at least one row, and a blank line counts as an (empty) row, exactly like
codegen/gen.py:155-166.  Coefficients are rationals (``1``, ``-.5``, ``1/2``)
or, for approximate (APA) algorithms, expressions in one parameter ``x``
(grammar of codegen/gen.py:101-147).  Values are kept exact: a coefficient
is a ``Coeff`` that is either a ``Fraction`` or a small polynomial/rational
expression tree in ``x``.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Sequence, Tuple

TABLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tables")


# ---------------------------------------------------------------------------
# Coefficients
# ---------------------------------------------------------------------------
class Coeff:
    """A coefficient: either an exact rational or an expression in ``x``.

    ``expr`` is a tuple tree: ('num', Fraction) | ('x',) | ('neg', e) |
    ('inv', e) | ('mul', e1, e2) | ('sum', [e...]).
    """

    __slots__ = ("expr", "text")

    def __init__(self, expr, text: str):
        self.expr = expr
        self.text = text

    # -- evaluation ---------------------------------------------------------
    def is_const(self) -> bool:
        return self.expr[0] == "num"

    @property
    def value(self) -> Fraction:
        if not self.is_const():
            raise ValueError("coefficient %r depends on x" % self.text)
        return self.expr[1]

    def is_zero(self) -> bool:
        return self.is_const() and self.expr[1] == 0

    def eval(self, x: float = 1e-8) -> float:
        return _eval(self.expr, x)

    def eval_exact(self, x: Fraction) -> Fraction:
        return _eval_exact(self.expr, x)

    def c_expr(self, scalar: str = "T") -> str:
        """C++ expression for the coefficient (``x`` is a double named x)."""
        return _c_expr(self.expr, scalar)

    def __repr__(self):
        return "Coeff(%s)" % self.text


def _eval(e, x):
    t = e[0]
    if t == "num":
        return float(e[1])
    if t == "x":
        return x
    if t == "neg":
        return -_eval(e[1], x)
    if t == "inv":
        return 1.0 / _eval(e[1], x)
    if t == "mul":
        return _eval(e[1], x) * _eval(e[2], x)
    if t == "sum":
        return sum(_eval(s, x) for s in e[1])
    raise ValueError(e)


def _eval_exact(e, x):
    t = e[0]
    if t == "num":
        return e[1]
    if t == "x":
        return x
    if t == "neg":
        return -_eval_exact(e[1], x)
    if t == "inv":
        return 1 / _eval_exact(e[1], x)
    if t == "mul":
        return _eval_exact(e[1], x) * _eval_exact(e[2], x)
    if t == "sum":
        return sum((_eval_exact(s, x) for s in e[1]), Fraction(0))
    raise ValueError(e)


def _c_expr(e, scalar):
    t = e[0]
    if t == "num":
        return repr(float(e[1]))
    if t == "x":
        return "x"
    if t == "neg":
        return "(-(%s))" % _c_expr(e[1], scalar)
    if t == "inv":
        return "(1.0 / (%s))" % _c_expr(e[1], scalar)
    if t == "mul":
        return "((%s) * (%s))" % (_c_expr(e[1], scalar), _c_expr(e[2], scalar))
    if t == "sum":
        return "(" + " + ".join(_c_expr(s, scalar) for s in e[1]) + ")"
    raise ValueError(e)


def parse_coeff(text: str) -> Coeff:
    """Parse one coefficient with the grammar of codegen/gen.py:101-147.

    ``*i`` is 1/(*), ``-*`` is -(*), ``*p`` (p a digit) is (*)^p, a single
    letter is the APA parameter x, ``(a+b)`` is a sum and ``c x...`` a
    numeric multiplier in front of an x-expression.
    """
    s = text.strip()
    try:
        return Coeff(("num", Fraction(s)), s)
    except (ValueError, ZeroDivisionError):
        pass
    return Coeff(_parse_expr(s), s)


def _parse_expr(s):
    if len(s) == 1:
        return ("x",)
    if s[0] == "(":
        if s[-1] != ")":
            raise ValueError("cannot parse coefficient %r" % s)
        return ("sum", [_parse_expr_or_num(p) for p in s[1:-1].split("+")])
    if s[0] == "-":
        return ("neg", _parse_expr_or_num(s[1:]))
    pos = s.find("x")
    if pos > 0:
        try:
            mult = Fraction(s[:pos])
            return ("mul", ("num", mult), _parse_expr(s[pos:]))
        except (ValueError, ZeroDivisionError):
            pass
    if s[-1] == "i":
        return ("inv", _parse_expr(s[:-1]))
    if s[-1].isdigit():
        base = _parse_expr(s[:-1])
        e = base
        for _ in range(int(s[-1]) - 1):
            e = ("mul", e, base)
        return e
    raise ValueError("cannot parse coefficient %r" % s)


def _parse_expr_or_num(s):
    s = s.strip()
    try:
        return ("num", Fraction(s))
    except (ValueError, ZeroDivisionError):
        return _parse_expr(s)


# ---------------------------------------------------------------------------
# Tables
# ---------------------------------------------------------------------------
@dataclass
class Table:
    """One fast algorithm <m,k,n;R>.

    U[p][r]: coefficient of A block p (p < m*k, row-major) or A_X temporary
    (p >= m*k) in S_r.  V likewise for B.  W[c][q]: coefficient of product q
    (q < R) or M_X temporary (q >= R) in C block c.  SX[i][p], TX[i][p]:
    definition of A_X(i+1), B_X(i+1) over blocks; MX[i][r]: definition of
    M_X(i+1) over products.
    """

    name: str
    m: int
    k: int
    n: int
    U: List[List[Coeff]]
    V: List[List[Coeff]]
    W: List[List[Coeff]]
    SX: List[List[Coeff]] = field(default_factory=list)
    TX: List[List[Coeff]] = field(default_factory=list)
    MX: List[List[Coeff]] = field(default_factory=list)
    path: str = ""

    @property
    def R(self) -> int:
        return len(self.U[0])

    @property
    def is_apa(self) -> bool:
        for grp in (self.U, self.V, self.W, self.SX, self.TX, self.MX):
            for row in grp:
                for c in row:
                    if not c.is_const():
                        return True
        return False

    def nnz(self) -> int:
        """Nonzeros over all groups (the number in the table's file name)."""
        tot = 0
        for grp in (self.U, self.V, self.W, self.SX, self.TX, self.MX):
            for row in grp:
                tot += sum(1 for c in row if not c.is_zero())
        return tot

    # column helpers -------------------------------------------------------
    def s_terms(self, r: int) -> List[Tuple[int, Coeff]]:
        """Nonzero (row, coeff) of U column r, in table row order."""
        return [(p, row[r]) for p, row in enumerate(self.U) if not row[r].is_zero()]

    def t_terms(self, r: int) -> List[Tuple[int, Coeff]]:
        return [(p, row[r]) for p, row in enumerate(self.V) if not row[r].is_zero()]

    def w_terms(self, c: int) -> List[Tuple[int, Coeff]]:
        return [(q, co) for q, co in enumerate(self.W[c]) if not co.is_zero()]

    def sx_terms(self, i: int) -> List[Tuple[int, Coeff]]:
        return [(p, co) for p, co in enumerate(self.SX[i]) if not co.is_zero()]

    def tx_terms(self, i: int) -> List[Tuple[int, Coeff]]:
        return [(p, co) for p, co in enumerate(self.TX[i]) if not co.is_zero()]

    def mx_terms(self, i: int) -> List[Tuple[int, Coeff]]:
        return [(q, co) for q, co in enumerate(self.MX[i]) if not co.is_zero()]

    # exact expansion ------------------------------------------------------
    def expanded(self, x: Optional[Fraction] = None):
        """Return (U', V', W') over blocks/products only, with the CSE
        temporaries substituted back in (exact Fractions)."""
        xv = Fraction(1, 10**8) if x is None else x

        def val(c):
            return c.eval_exact(xv)

        mk, kn, mn, R = self.m * self.k, self.k * self.n, self.m * self.n, self.R
        Ue = [[val(self.U[p][r]) for r in range(R)] for p in range(mk)]
        for i, row in enumerate(self.SX):
            for r in range(R):
                cx = val(self.U[mk + i][r])
                if cx:
                    for p in range(mk):
                        Ue[p][r] += cx * val(row[p])
        Ve = [[val(self.V[p][r]) for r in range(R)] for p in range(kn)]
        for i, row in enumerate(self.TX):
            for r in range(R):
                cx = val(self.V[kn + i][r])
                if cx:
                    for p in range(kn):
                        Ve[p][r] += cx * val(row[p])
        We = [[val(self.W[c][q]) for q in range(R)] for c in range(mn)]
        for i, row in enumerate(self.MX):
            for c in range(mn):
                cx = val(self.W[c][R + i])
                if cx:
                    for q in range(R):
                        We[c][q] += cx * val(row[q])
        return Ue, Ve, We


def read_table_groups(path: str) -> List[List[List[str]]]:
    """Split a table file into groups of rows of coefficient strings.

    Mirrors codegen/gen.py:150-170: '#' closes the current group only when it
    holds at least one row; a blank line is an empty row (so an empty
    substitution group written as a blank line still occupies its slot)."""
    groups: List[List[List[str]]] = []
    cur: List[List[str]] = []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                if cur:
                    groups.append(cur)
                    cur = []
            else:
                cur.append(line.split())
    groups.append(cur)
    return groups


def load_table(path: str, dims: Optional[Tuple[int, int, int]] = None, name: str = "") -> Table:
    groups = read_table_groups(path)
    if len(groups) < 3:
        raise ValueError("%s: expected three sets of coefficients" % path)

    def conv(grp):
        return [[parse_coeff(v) for v in row] for row in grp if row]

    U, V, W = conv(groups[0]), conv(groups[1]), conv(groups[2])
    SX = conv(groups[3]) if len(groups) > 3 else []
    TX = conv(groups[4]) if len(groups) > 4 else []
    MX = conv(groups[5]) if len(groups) > 5 else []
    R = len(U[0])
    mk = len(U) - len(SX)
    kn = len(V) - len(TX)
    mn = len(W)
    if dims is None:
        m2 = Fraction(mk * mn, kn)
        m = int(round(float(m2) ** 0.5))
        if m <= 0 or m * m != m2:
            raise ValueError("%s: cannot infer dims" % path)
        k, n = mk // m, mn // m
        dims = (m, k, n)
    m, k, n = dims
    if (m * k, k * n, m * n) != (mk, kn, mn):
        raise ValueError("%s: dims %s do not match table shape" % (path, dims))
    for grp, width, what in ((U, R, "U"), (V, R, "V"), (W, R + len(MX), "W"),
                             (SX, m * k, "SX"), (TX, k * n, "TX"), (MX, R, "MX")):
        for row in grp:
            if len(row) != width:
                raise ValueError("%s: %s row has %d entries, expected %d" % (path, what, len(row), width))
    return Table(name=name or os.path.basename(path), m=m, k=k, n=n, U=U, V=V, W=W,
                 SX=SX, TX=TX, MX=MX, path=path)


def brent_residual(t: Table, x: Optional[Fraction] = None) -> Fraction:
    """Max |sum_r U'[(i,j),r] V'[(j',l),r] W'[(i',l'),r] - delta| over all
    index combinations (the Brent equations; codegen/verify.py:110-132).
    Zero for an exact algorithm; O(x) for an APA one."""
    Ue, Ve, We = t.expanded(x)
    m, k, n, R = t.m, t.k, t.n, t.R
    worst = Fraction(0)
    for i in range(m):
        for j in range(k):
            a = i * k + j
            for j2 in range(k):
                for l in range(n):
                    b = j2 * n + l
                    # product coefficient vector for this (a, b) pair
                    uv = [Ue[a][r] * Ve[b][r] for r in range(R)]
                    for i2 in range(m):
                        for l2 in range(n):
                            c = i2 * n + l2
                            s = sum((uv[r] * We[c][r] for r in range(R)), Fraction(0))
                            want = 1 if (i == i2 and j == j2 and l == l2) else 0
                            d = abs(s - want)
                            if d > worst:
                                worst = d
    return worst


# ---------------------------------------------------------------------------
# Registry: committed namespace name -> (table file, dims)
# The namespace names are the ones of the reference's generated headers
# (algorithms/*.hpp, derived from the table file name by
# codegen/gen.py:902-906); the enum-level dispatch is
# algorithms/all_algorithms.hpp:93-243.
# ---------------------------------------------------------------------------
_REGISTRY: Dict[str, Tuple[str, Tuple[int, int, int]]] = {
    "strassen": ("strassen", (2, 2, 2)),
    "classical222_8_24": ("classical222-8-24", (2, 2, 2)),
    "classical333_27_81": ("classical333-27-81", (3, 3, 3)),
    "classical423_24_72": ("classical423-24-72", (4, 2, 3)),
    "bini322_10_52_approx": ("bini322-10-52-approx", (3, 2, 2)),
    "schonhage333_21_117_approx": ("schonhage333-21-117-approx", (3, 3, 3)),
    "smirnov333_20_182_approx": ("smirnov333-20-182-approx", (3, 3, 3)),
    "grey234_20_144": ("grey234-20-144", (2, 3, 4)),
    "grey243_20_144": ("grey243-20-144", (2, 4, 3)),
    "grey252_18_99": ("grey252-18-99", (2, 5, 2)),
    "grey322_11_50": ("grey322-11-50", (3, 2, 2)),
    "grey323_15_103": ("grey323-15-103", (3, 2, 3)),
    "grey323_15_89": ("eliminated/grey323-15-89", (3, 2, 3)),
    "grey324_20_144": ("grey324-20-144", (3, 2, 4)),
    "grey332_15_103": ("grey332-15-103", (3, 3, 2)),
    "grey333_23_125": ("eliminated/grey333-23-125", (3, 3, 3)),
    "grey333_23_152": ("grey333-23-152", (3, 3, 3)),
    "grey333_23_221": ("grey333-23-221", (3, 3, 3)),
    "grey342_20_144": ("grey342-20-144", (3, 4, 2)),
    "grey343_29_234": ("grey343-29-234", (3, 4, 3)),
    "grey422_14_84": ("grey422-14-84", (4, 2, 2)),
    "fast423_130": ("fast423-130", (4, 2, 3)),
    "fast423_134": ("fast423-134", (4, 2, 3)),
    "fast423_138": ("fast423-138", (4, 2, 3)),
    "fast423_156": ("fast423-156", (4, 2, 3)),
    "grey423_20_144": ("grey423-20-144", (4, 2, 3)),
    "grey424_26_206": ("eliminated/grey424-26-206", (4, 2, 4)),
    "grey424_26_257": ("grey424-26-257", (4, 2, 4)),
    "grey432_20_144": ("grey432-20-144", (4, 3, 2)),
    "grey433_29_234": ("grey433-29-234", (4, 3, 3)),
    "grey442_26_257": ("grey442-26-257", (4, 4, 2)),
    "grey522_18_99": ("grey522-18-99", (5, 2, 2)),
    "hk323_15_84": ("eliminated/hk323-15-84", (3, 2, 3)),
    "hk323_15_94": ("hk323-15-94", (3, 2, 3)),
    "hk332_15_94": ("hk332-15-94", (3, 3, 2)),
    "smirnov333_23_128": ("eliminated/smirnov333-23-128", (3, 3, 3)),
    "smirnov333_23_139": ("smirnov333-23-139", (3, 3, 3)),
    "smirnov336_40_960": ("smirnov336-40-960", (3, 3, 6)),
    "smirnov363_40_960": ("smirnov363-40-960", (3, 6, 3)),
    "smirnov633_40_960": ("smirnov633-40-960", (6, 3, 3)),
    "tichavsky_kovac336_40_960": ("tichavsky_kovac336-40-960", (3, 3, 6)),
    "tichavsky_kovac363_40_960": ("tichavsky_kovac363-40-960", (3, 6, 3)),
}

# Enum names of algorithms/all_algorithms.hpp:52-90 -> namespace, including
# the reference's dispatch of FAST323_15_89 to grey323_15_103
# (algorithms/all_algorithms.hpp:159-160) and MKL -> classical222 at depth 0
# (algorithms/all_algorithms.hpp:98-101).
ENUM_TO_NAMESPACE: Dict[str, str] = {
    "MKL": "classical222_8_24",
    "BINI322_10_52_APPROX": "bini322_10_52_approx",
    "CLASSICAL222": "classical222_8_24",
    "CLASSICAL423": "classical423_24_72",
    "CLASSICAL333": "classical333_27_81",
    "SMIRNOV333_23_128": "smirnov333_23_128",
    "SMIRNOV333_23_139": "smirnov333_23_139",
    "SMIRNOV336_40_960": "smirnov336_40_960",
    "SMIRNOV363_40_960": "smirnov363_40_960",
    "SMIRNOV633_40_960": "smirnov633_40_960",
    "HK332_15_94": "hk332_15_94",
    "HK323_15_94": "hk323_15_94",
    "HK323_15_84": "hk323_15_84",
    "STRASSEN": "strassen",
    "FAST322_11_50": "grey322_11_50",
    "FAST323_15_103": "grey323_15_103",
    "FAST323_15_89": "grey323_15_103",
    "FAST332_15_103": "grey332_15_103",
    "FAST333_23_125": "grey333_23_125",
    "FAST333_23_152": "grey333_23_152",
    "FAST333_23_221": "grey333_23_221",
    "FAST432_20_144": "grey432_20_144",
    "FAST423_20_144": "grey423_20_144",
    "FAST324_20_144": "grey324_20_144",
    "FAST342_20_144": "grey342_20_144",
    "FAST234_20_144": "grey234_20_144",
    "FAST243_20_144": "grey243_20_144",
    "FAST422_14_84": "grey422_14_84",
    "FAST442_26_257": "grey442_26_257",
    "FAST424_26_257": "grey424_26_257",
    "FAST424_26_206": "grey424_26_206",
    "FAST433_29_234": "grey433_29_234",
    "FAST343_29_234": "grey343_29_234",
    "FAST522_18_99": "grey522_18_99",
    "SCHONHAGE333_21_117_APPROX": "schonhage333_21_117_approx",
}


# ---------------------------------------------------------------------------
# Permuted algorithms (codegen/convert.py:61-99,136-156).  From <m,k,n;R>:
#   transposition  C^T = B^T A^T  gives <n,k,m;R>:
#       U'[(j,i)] = V[(i,j)],  V'[(j,i)] = U[(i,j)],  W'[(j,i)] = W[(i,j)];
#   the cyclic symmetry of the multiplication tensor gives <n,m,k;R>:
#       U'[(j,i)] = W[(i,j)],  V' = U,  W'[(j,i)] = V[(i,j)];
# and compositions of the two give all six orders.  Coefficients are
# unchanged; only the block numbering moves, so a permuted table is exact
# iff its source is (checked by the Brent equations in tests/test_tables.py).
# Tables with CSE groups are not permuted (their temporaries would change
# sides); the registry derives only permutations absent from the committed
# tables.
# ---------------------------------------------------------------------------
def _transposed_rows(rows, r: int, c: int):
    """Rows of an r x c block grid (row-major) reordered to the c x r grid."""
    if len(rows) != r * c:
        raise ValueError("coefficient group has %d rows, expected %d" % (len(rows), r * c))
    return [list(rows[i * c + j]) for j in range(c) for i in range(r)]


def _swap(t: Table) -> Table:
    m, k, n = t.m, t.k, t.n
    return Table(name=t.name, m=n, k=k, n=m, U=_transposed_rows(t.V, k, n),
                 V=_transposed_rows(t.U, m, k), W=_transposed_rows(t.W, m, n), path=t.path)


def _cyclic(t: Table) -> Table:
    m, k, n = t.m, t.k, t.n
    return Table(name=t.name, m=n, k=m, n=k, U=_transposed_rows(t.W, m, n),
                 V=[list(row) for row in t.U], W=_transposed_rows(t.V, k, n), path=t.path)


def permute_table(t: Table, perm: Tuple[int, int, int], name: str = "") -> Table:
    """The <d[p0], d[p1], d[p2]> algorithm from t = <d0, d1, d2> (perm as in
    codegen/convert.py's s1,s2,s3 argument)."""
    if t.SX or t.TX or t.MX:
        raise ValueError("%s: tables with common-subexpression groups are not permuted" % t.name)
    perm = tuple(perm)
    if perm == (0, 1, 2):
        out = Table(name=t.name, m=t.m, k=t.k, n=t.n, U=[list(r) for r in t.U],
                    V=[list(r) for r in t.V], W=[list(r) for r in t.W], path=t.path)
    elif perm == (2, 0, 1):
        out = _cyclic(t)
    elif perm == (1, 2, 0):
        out = _cyclic(_cyclic(t))
    elif perm == (2, 1, 0):
        out = _swap(t)
    elif perm == (0, 2, 1):
        out = _swap(_cyclic(_cyclic(t)))
    elif perm == (1, 0, 2):
        out = _swap(_cyclic(t))
    else:
        raise ValueError("invalid permutation %r" % (perm,))
    dims = (t.m, t.k, t.n)
    assert (out.m, out.k, out.n) == (dims[perm[0]], dims[perm[1]], dims[perm[2]])
    out.name = name or t.name
    return out


# derived name -> (source namespace, permutation)
_DERIVED: Dict[str, Tuple[str, Tuple[int, int, int]]] = {
    "grey232_11_50": ("grey322_11_50", (1, 0, 2)),
    "grey223_11_50": ("grey322_11_50", (1, 2, 0)),
    "grey242_14_84": ("grey422_14_84", (1, 0, 2)),
    "grey224_14_84": ("grey422_14_84", (1, 2, 0)),
    "grey225_18_99": ("grey522_18_99", (1, 2, 0)),
    "grey244_26_257": ("grey424_26_257", (1, 0, 2)),
    "grey334_29_234": ("grey343_29_234", (0, 2, 1)),
    "grey233_15_103": ("grey323_15_103", (1, 0, 2)),
    "hk233_15_94": ("hk323_15_94", (1, 0, 2)),
    "tichavsky_kovac633_40_960": ("tichavsky_kovac336_40_960", (2, 1, 0)),
}


def derived_algorithms() -> Dict[str, Tuple[str, Tuple[int, int, int]]]:
    return dict(_DERIVED)


def algorithm_names() -> List[str]:
    return sorted(list(_REGISTRY) + list(_DERIVED))


def get_table(name: str) -> Table:
    if name in _DERIVED:
        src, perm = _DERIVED[name]
        return permute_table(get_table(src), perm, name)
    if name not in _REGISTRY:
        raise KeyError("unknown algorithm %r" % name)
    rel, dims = _REGISTRY[name]
    return load_table(os.path.join(TABLE_DIR, rel), dims, name)


def default_x(enum_name: str, num_steps: int) -> float:
    """APA parameter chosen by RunAlgorithm (algorithms/all_algorithms.hpp:102-121
    for Bini; the Schonhage branch follows it)."""
    if enum_name == "BINI322_10_52_APPROX":
        return {2: 1e-4, 3: 1e-2, 4: 1e-1}.get(num_steps, 1e-8)
    if enum_name == "SCHONHAGE333_21_117_APPROX":
        return {1: 1e-4, 2: 1e-2, 3: 1e-1, 4: 1e-1}.get(num_steps, 1e-4)
    return 1e-8
