"""Write tests/golden/d3q15_tables.json from the reference's table source.

Reads lbswim/liblbswim/d3q15.c of the reference checkout as TEXT and
evaluates each right-hand side, which is a literal such as ``-1./3.`` or
``1./72.``, to the IEEE double that C's constant folding produces (Python
float division is correctly rounded, like C's).  Nothing from the reference is
imported or executed.  The JSON is the golden fixture that pins the D3Q15
tables of the oracle and of the HIP library (d3q15.c:2-186).

Usage:  python tests/golden/make_d3q15_tables.py [/root/reference]
"""
import json
import os
import re
import sys
from fractions import Fraction

Q, D = 15, 3
LINE = re.compile(r"lat->(\w+)((?:\[\d+\])*)\s*=\s*([-+0-9./ ]+);")


def _literal(text):
    """Evaluate 'a./b.' or 'n' the way a C compiler folds it."""
    text = text.replace(" ", "")
    if "/" in text:
        num, den = text.split("/")
        return float(Fraction(num.rstrip(".") or "0") / Fraction(den.rstrip(".")))
    return float(Fraction(text.rstrip(".")))


def main(ref_root):
    src = os.path.join(ref_root, "lbswim", "liblbswim", "d3q15.c")
    with open(src) as fh:
        body = fh.read()
    tables = {
        "cs2": None,
        "w": [None] * Q,
        "xi": [[None] * D for _ in range(Q)],
        "ci": [[None] * D for _ in range(Q)],
        "complement": [None] * Q,
        "Q": [[[None] * D for _ in range(D)] for _ in range(Q)],
        "norms": [None] * Q,
        "mm": [[None] * Q for _ in range(Q)],
        "mmi": [[None] * Q for _ in range(Q)],
    }
    for name, idx, rhs in LINE.findall(body):
        val = _literal(rhs)
        ids = [int(i) for i in re.findall(r"\[(\d+)\]", idx)]
        if name in ("ci", "complement"):
            val = int(val)
        if not ids:
            tables[name] = val
            continue
        slot = tables[name]
        for i in ids[:-1]:
            slot = slot[i]
        slot[ids[-1]] = val

    def _check(x):
        if isinstance(x, list):
            for y in x:
                _check(y)
        elif x is None:
            raise SystemExit("incomplete table in %s" % src)

    _check(list(tables.values()))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "d3q15_tables.json")
    with open(out, "w") as fh:
        json.dump({"source": "lbswim/liblbswim/d3q15.c:2-186", "tables": tables}, fh, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
