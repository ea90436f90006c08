"""Writes hnc3d_regression.json: the numbers of the reference's HNC3D
regression outputs test/out/*.hnc3d.g1 and water.hnc3d.g2 (made by
test/Makefile with HNC-FLAGS: --N 96 --L 10 --beta 1.6889 --rho
0.033427745 --norm-tol 1e-12 --closure KH --solvent water --solvent-3d).

Data only: per file the moments of 1 - g (charge, dipole, second-moment
matrix, as the strings the reference prints) and the print_info / mu lines
test/summary.sh extracts.  Run in a tree where the reference is at
/root/reference:  python tests/golden/make_hnc3d_golden.py
"""
import json
import os
import re

SRC = "/root/reference/test/out"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hnc3d_regression.json")
SOLUTES = ["water", "hydrogen chloride", "carbon disulfide", "methanol", "butanoic acid",
           "hexane"]


def parse(path):
    text = open(path).read()
    moments = []
    for block in re.split(r";;; Moments from \S+\n", text)[1:]:
        nums = re.findall(r'"(-?[0-9.]+)"', block.split("\n;;;")[0])[:13]
        moments.append({"q": nums[0], "d": nums[1:4], "Q": [nums[4:7], nums[7:10], nums[10:13]]})
    info = {}
    num = r"(-?[0-9]+\.[0-9]+)"
    pats = {"q_v": rf"q_v = {num}", "q_u": rf"q_u = {num}",
            "U_v.rho_N": rf"<U_v\|ρ_N> = {num}", "U_v.rho_u": rf"<U_v\|ρ_u> = {num}",
            "rho_v.U_u": rf"<ρ_v\|U_u> = {num}",
            "mu_hnc": rf"mu = {num} kcal \(HNC\)", "mu_kh": rf"mu = {num} kcal \(KH\)",
            "mu_gf": rf"mu = {num} kcal \(GF\)"}
    for k, p in pats.items():
        m = re.search(p, text)
        if m:
            info[k] = m.group(1)
    for k, tag in (("d_n", "ρ_n"), ("d_u", "ρ_u"), ("d_v", "ρ_v")):
        m = re.search(rf"\|<x\|{tag}>\| = \|\s*{num},\s*{num},\s*{num}\|", text)
        if m:
            info[k] = [m.group(1), m.group(2), m.group(3)]
    return {"moments": moments, "info": info}


def main():
    out = {"solvent": parse(os.path.join(SRC, "water.hnc3d.g2")), "solutes": {}}
    for s in SOLUTES:
        out["solutes"][s] = parse(os.path.join(SRC, s.replace(" ", "_") + ".hnc3d.g1"))
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, ensure_ascii=False)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
