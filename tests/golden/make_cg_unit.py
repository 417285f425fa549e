"""Extract the reference's conjugate-gradients unit-test data into
cg_unit.json: the 14-node / 24-tet mesh of
tests/unit/LinearSolver/TestConjugateGradients.cpp and the expected norm of
b and final residual of its 1-DOF and 3-DOF serial tests (tolerance 1e-12),
and the two-partition ("2PE") tests: per partition the node coordinates,
tets (local ids), psup, local-to-global node map and node communication map,
with the expected norms.  The file is parsed as text (the mesh arrays and
the literal goldens)."""
import json
import os
import re

SRC = "/root/reference/tests/unit/LinearSolver/TestConjugateGradients.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cg_unit.json")


def _ints(t):
    return [int(v) for v in re.findall(r"\d+", t)]


def _block(seg, name):
    """The brace-balanced initializer that follows `name`."""
    i = seg.index(name)
    i = seg.index("{", i)
    depth, j = 0, i
    while True:
        if seg[j] == "{":
            depth += 1
        elif seg[j] == "}":
            depth -= 1
            if depth == 0:
                return seg[i:j + 1]
        j += 1


def _groups(t, level):
    """The brace groups at nesting depth `level` of an initializer."""
    out, depth, start = [], 0, None
    for j, ch in enumerate(t):
        if ch == "{":
            depth += 1
            if depth == level:
                start = j
        elif ch == "}":
            if depth == level:
                out.append(t[start:j + 1])
            depth -= 1
    return out


def two_pe(seg):
    """Per-partition data of a 2PE test body."""
    psup = [[_ints(g) for g in _groups(p, 2)] for p in _groups(_block(seg, "psup"), 2)]
    crd = _block(seg, "coord")
    coord = [[[float(v) for v in re.findall(r"-?[0-9.]+", c)] for c in _groups(p, 4)]
             for p in _groups(crd, 3)]
    inpoel = [_ints(g) for g in _groups(_block(seg, "inpoel"), 2)]
    gid = [_ints(g) for g in _groups(_block(seg, "> gid"), 2)]
    comm = re.findall(r"nodecommap\.back\(\)\[(\d+)\]\.insert\( \{([^}]*)\}", seg)
    nodecommap = [{int(c): _ints(n)} for c, n in comm]
    return {"psup": psup, "coord": coord, "inpoel": inpoel, "gid": gid,
            "nodecommap": nodecommap,
            "maxit": int(re.search(r"maxit = \*/ (\d+)", seg).group(1)),
            "tol": float(re.search(r"tol = \* \*/ ([0-9.e+-]+)", seg).group(1)),
            "normb": float(re.search(r"normb_ex = \*/ ([0-9.e+-]+)", seg).group(1)),
            "normres": float(re.search(r"normres_ex = \*/ ([0-9.e+-]+)", seg).group(1))}


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
    two = {}
    for k, tname, nxt in ((1, "test< 2 >", "test< 3 >"), (3, "test< 4 >", "#if defined")):
        seg = txt[txt.index(tname):]
        seg = seg[:seg.index(nxt)]
        two[k] = two_pe(seg)
    json.dump({"source": "tests/unit/LinearSolver/TestConjugateGradients.cpp",
               "inpoel_1based": inpoel, "coord": coord, "ncomp": gold,
               "two_pe": two, "tolerance": 1e-12}, open(OUT, "w"), indent=1)
    print("wrote", OUT, gold)


if __name__ == "__main__":
    main()
