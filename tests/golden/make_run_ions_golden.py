"""Writes run_ions_out.json: the mu table of the reference's
test/out/run-ions-out (made by test/run-ions.scm with its *settings*:
cSPC/E water, rmax 160, nrad 4096, --L 10 --N 64, rho 0.0333295,
beta 1.6889, dielectric 78.4, one solvent chi per closure).

Data only: {"1d": {closure: {ion: "mu"}}, "3d": {...}} with the strings the
reference printed (4 decimals).  Run in a tree where the reference is at
/root/reference:  python tests/golden/make_run_ions_golden.py
"""
import json
import os
import re

SRC = "/root/reference/test/out/run-ions-out"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "run_ions_out.json")


def parse(text):
    out = {}
    for blk in re.finditer(r'\((KH|PSE\d|HNC) ((?:\("[^"]+" "[^"]+"\)\s*)+)\)', text):
        out[blk.group(1)] = dict(re.findall(r'\("([^"]+)" "([^"]+)"\)', blk.group(2)))
    return out


def main():
    text = open(SRC).read()
    i3 = text.index("RISM-3D")
    out = {"1d": parse(text[:i3]), "3d": parse(text[i3:])}
    assert all(len(v) == 9 for d in out.values() for v in d.values())
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
