"""Write d3q19_tables.json: the reference's lattice tables as data.

Reads Source1.f90 of the reference checkout as text (it is never imported or
run here) and extracts the MRT matrix DATA statement (S:105-123), the D3Q19
velocity, opposite and mirror tables (S:132-138) and the weights (S:140-142).
Run from the repository root:  python tests/golden/make_tables_fixture.py
"""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Source1.f90"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "d3q19_tables.json")


def main():
    text = open(SRC, "rb").read().decode("latin-1").replace("\r", "")
    m = re.search(r"data\s+M\s*/(.*?)/", text, re.S | re.I)
    vals = [int(v) for v in re.findall(r"-?\d+", m.group(1).replace("&", " "))]
    assert len(vals) == 361, len(vals)
    # Fortran DATA fills column-major: the first 19 values are M(0:18, 0).
    M = [[vals[j * 19 + k] for j in range(19)] for k in range(19)]

    def intvec(name):
        mm = re.search(r"^\s*%s\s*=\s*\(/(.*?)/\)" % name, text, re.M | re.I)
        return [int(v) for v in mm.group(1).split(",")]

    tables = {name: intvec(name) for name in ("cix", "ciy", "ciz", "opc", "symx", "symy", "symz")}
    w = {}
    for mm in re.finditer(r"^\s*wi\((\d+)(?::(\d+))?\)\s*=\s*1\.0d0/(\d+)\.0d0", text, re.M | re.I):
        lo = int(mm.group(1))
        hi = int(mm.group(2) or lo)
        for i in range(lo, hi + 1):
            w[i] = int(mm.group(3))
    wi_den = [w[i] for i in range(19)]
    out = {"source": "Source1.f90 S:105-123 (M), S:132-138, S:140-142",
           "M": M, "wi_denominator": wi_den, **tables}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
