#!/usr/bin/env python3
"""Write header_structure.json: the add/product structure of the reference's
committed generated headers (algorithms/*.hpp), as data.

For every algorithm it records, from the code the reference actually ships:
  * per product r: the S operand (blocks/temporaries with coefficients, from
    the S_Add<r> body and its call site, or the bare block passed to
    FastMatmulRecursive), the same for T, and the product of the
    UpdateMultiplier coefficients applied to M_r;
  * per C block: the M terms and coefficients of its M_Add body;
  * the CSE temporaries A_X/B_X/M_X (SX_/TX_/MX_Add bodies and call sites).
Only these numbers are written; no reference text is kept.

Runs only where the reference tree exists (this container):
    python tests/golden/extract_header_structure.py [/root/reference]
"""
from __future__ import annotations

import json
import os
import re
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))

HEADERS = {
    "strassen": "strassen.hpp", "classical222_8_24": "classical222.hpp",
    "classical333_27_81": "classical333.hpp", "classical423_24_72": "classical423.hpp",
    "grey234_20_144": "fast234_20_144.hpp", "grey243_20_144": "fast243_20_144.hpp",
    "grey252_18_99": "fast252_18_99.hpp", "grey322_11_50": "fast322_11_50.hpp",
    "grey323_15_103": "fast323_15_103.hpp", "grey323_15_89": "fast323_15_89.hpp",
    "grey324_20_144": "fast324_20_144.hpp", "grey332_15_103": "fast332_15_103.hpp",
    "grey333_23_125": "fast333_23_125.hpp", "grey333_23_152": "fast333_23_152.hpp",
    "grey333_23_221": "fast333_23_221.hpp", "grey342_20_144": "fast342_20_144.hpp",
    "grey343_29_234": "fast343_29_234.hpp", "grey422_14_84": "fast422_14_84.hpp",
    "fast423_130": "fast423_20_130.hpp", "fast423_134": "fast423_20_134.hpp",
    "fast423_138": "fast423_20_138.hpp", "fast423_156": "fast423_20_156.hpp",
    "grey423_20_144": "fast423_20_144.hpp", "grey424_26_206": "fast424_26_206.hpp",
    "grey424_26_257": "fast424_26_257.hpp", "grey432_20_144": "fast432_20_144.hpp",
    "grey433_29_234": "fast433_29_234.hpp", "grey442_26_257": "fast442_26_257.hpp",
    "grey522_18_99": "fast522_18_99.hpp", "hk323_15_84": "hk323_15_84.hpp",
    "hk323_15_94": "hk323_15_94.hpp", "hk332_15_94": "hk332_15_94.hpp",
    "smirnov333_23_128": "smirnov333_23_128.hpp", "smirnov333_23_139": "smirnov333_23_139.hpp",
    "smirnov336_40_960": "smirnov336_40_960.hpp", "smirnov363_40_960": "smirnov363_40_960.hpp",
    "smirnov633_40_960": "smirnov633_40_960.hpp",
    "tichavsky_kovac336_40_960": "tichavsky_kovac336_40_960.hpp",
    "tichavsky_kovac363_40_960": "tichavsky_kovac363_40_960.hpp",
}

FUNC_RE = re.compile(r"void (S|T|M|SX|TX|MX)_Add(\d+)\(([^)]*)\)\s*\{")
TERM_RE = re.compile(r"(-?)\s*(?:Scalar\(([^)]*)\)\s*\*\s*)?data(\w+?)\[i \+ j \* stride\w+\]")
CALL_RE = re.compile(r"\b(S|T|M|SX|TX|MX)_Add(\d+)\(([^;]*)\);")
REC_RE = re.compile(r"FastMatmulRecursive\(locker, mem_mngr, (\w+), (\w+), (\w+),")
UPD_RE = re.compile(r"\b(M\d+)\.UpdateMultiplier\(Scalar\(([^)]*)\)\)")


def frac(text):
    return Fraction(text.strip())


def parse_bodies(src):
    """kind -> index -> list of (param_position, coeff) in term order."""
    out = {}
    for mt in FUNC_RE.finditer(src):
        kind, idx, params = mt.group(1), int(mt.group(2)), mt.group(3)
        pnames = [p.split("&")[-1].strip() for p in params.split(",")]
        body_start = mt.end()
        line_end = src.index(";", src.index("dataC[i + j * strideC] =", body_start))
        expr = src[src.index("=", src.index("dataC[i + j * strideC] =", body_start)) + 1:line_end]
        expr = expr.split("+ beta *")[0]
        terms = []
        for t in TERM_RE.finditer(expr):
            sign, coef, name = t.group(1), t.group(2), t.group(3)
            c = frac(coef) if coef else Fraction(1)
            if sign == "-":
                c = -c
            terms.append((pnames.index(name), c))
        out.setdefault(kind, {})[idx] = terms
    return out


def extract(path):
    src = open(path).read()
    rec_start = src.index("void FastMatmulRecursive(")
    rec = src[rec_start:]
    bodies = parse_bodies(src[:rec_start] + src[src.index("}", rec_start):] if False else src)
    calls = {}
    for c in CALL_RE.finditer(rec):
        kind, idx = c.group(1), int(c.group(2))
        args = [a.strip() for a in c.group(3).split(",")]
        calls.setdefault(kind, {})[idx] = args
    res = {"products": [], "C": {}, "SX": {}, "TX": {}, "MX": {}}

    def combo(kind, idx):
        args = calls[kind][idx]
        return [[args[p], str(c)] for p, c in bodies[kind][idx]], args

    for kind, key in (("SX", "SX"), ("TX", "TX"), ("MX", "MX")):
        for idx in sorted(calls.get(kind, {})):
            terms, args = combo(kind, idx)
            res[key][args[len(terms)]] = terms
    # products in call order, with the multiplier updates since the last call
    pos = 0
    for r in REC_RE.finditer(rec):
        if r.group(3) == "C":
            continue
        seg = rec[pos:r.start()]
        pos = r.end()
        a, b, mname = r.group(1), r.group(2), r.group(3)
        mult = Fraction(1)
        for u in UPD_RE.finditer(seg):
            if u.group(1) == mname:
                mult *= frac(u.group(2))
        prod = {"M": mname, "mult": str(mult)}
        for side, opnd, kind in (("S", a, "S"), ("T", b, "T")):
            m = re.fullmatch(r"[ST](\d+)", opnd)
            if m and int(m.group(1)) in calls.get(kind, {}):
                prod[side] = combo(kind, int(m.group(1)))[0]
            else:
                prod[side] = [[opnd, "1"]]
        res["products"].append(prod)
    for idx in sorted(calls.get("M", {})):
        terms, args = combo("M", idx)
        res["C"][args[len(terms)]] = terms
    return res


def main(argv):
    ref = argv[1] if len(argv) > 1 else "/root/reference"
    alg_dir = os.path.join(ref, "algorithms")
    out = {}
    for name, hdr in sorted(HEADERS.items()):
        out[name] = extract(os.path.join(alg_dir, hdr))
    path = os.path.join(HERE, "header_structure.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote", path, len(out), "algorithms")


if __name__ == "__main__":
    main(sys.argv)
