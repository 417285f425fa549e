"""Extract the reference's conjugate-gradients unit-test data into
cg_unit.json: the 14-node / 24-tet mesh of
tests/unit/LinearSolver/TestConjugateGradients.cpp and the expected norm of
b and final residual of its 1-DOF and 3-DOF serial tests (tolerance 1e-12).
The file is parsed as text (the mesh arrays and the literal goldens)."""
import json
import os
import re

SRC = "/root/reference/tests/unit/LinearSolver/TestConjugateGradients.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cg_unit.json")


def main():
    txt = open(SRC).read()
    t1 = txt[txt.index("test< 1 >"):txt.index("test< 2 >")]
    inp = re.search(r"inpoel \{(.*?)\};", t1, re.S).group(1)
    inpoel = [int(v) for v in re.findall(r"\d+", inp)]
    crd = re.search(r"coord \{\{(.*?)\}\};", t1, re.S).group(1)
    rows = re.findall(r"\{\{(.*?)\}\}", crd, re.S)
    coord = [[float(v) for v in r.split(",")] for r in rows]
    gold = {}
    for k, tname in ((1, "test< 1 >"), (3, "test< 3 >")):
        seg = txt[txt.index(tname):]
        seg = seg[:seg.index("cg[0].setup")]
        gold[k] = {
            "maxit": int(re.search(r"maxit = \*/ (\d+)", seg).group(1)),
            "tol": float(re.search(r"tol = \* \*/ ([0-9.e+-]+)", seg).group(1)),
            "normb": float(re.search(r"normb_ex = \*/ ([0-9.e+-]+)", seg).group(1)),
            "normres": float(re.search(r"normres_ex = \*/ ([0-9.e+-]+)", seg).group(1)),
        }
    json.dump({"source": "tests/unit/LinearSolver/TestConjugateGradients.cpp",
               "inpoel_1based": inpoel, "coord": coord, "ncomp": gold,
               "tolerance": 1e-12}, open(OUT, "w"), indent=1)
    print("wrote", OUT, gold)


if __name__ == "__main__":
    main()
