"""Writes rta_bulk_poly.json: the constants of the relaxation-time bulk delta-f
polynomials (bulk_deltaf_kind 1) as they stand in the reference's source,
FSSW::getbulkvisCoefficients(T) (src/FSSW.cpp:1109-1136).  The two
coefficients are 10th-order polynomials in T [1/fm]; each term is read with
its sign and power from the source text, so the fixture is the reference's
data, not a retyped copy of the oracle's table.

Run from the repository root: python tests/golden/make_deltaf_golden.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/FSSW.cpp"

TERM = re.compile(r"([+-]?)\s*([0-9]+\.?[0-9]*(?:[eE][+-]?[0-9]+)?)\s*(?:\*\s*Tdec_fm_power\[(\d+)\])?")


def _poly(expr):
    coef = [0.0] * 11
    seen = set()
    for sign, num, power in TERM.findall(expr):
        k = int(power) if power else 0
        if k in seen:
            raise ValueError(f"power {k} twice in {expr!r}")
        seen.add(k)
        coef[k] = -float(num) if sign == "-" else float(num)
    if seen != set(range(11)):
        raise ValueError(f"expected powers 0..10, got {sorted(seen)}")
    return coef


def main():
    text = open(SRC).read()
    fn = text.index("void FSSW::getbulkvisCoefficients(\n            const double Tdec,")
    start = text.index("bulk_deltaf_kind_ == 1", fn)
    stop = text.index("bulk_deltaf_kind_ == 2", start)
    block = text[start:stop]
    out = {}
    for i in (0, 1):
        m = re.search(r"bulkvisCoefficients\[%d\]\s*=\s*\((.*?)\);" % i, block, re.S)
        out[f"c{i}"] = _poly(m.group(1))
    out["source"] = "src/FSSW.cpp, FSSW::getbulkvisCoefficients(T), bulk_deltaf_kind_ == 1"
    out["variable"] = "T / hbarC [1/fm]"
    with open(os.path.join(HERE, "rta_bulk_poly.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "rta_bulk_poly.json"))


if __name__ == "__main__":
    main()
