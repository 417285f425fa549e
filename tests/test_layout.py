"""CPU emulation of the rhs block layout (tests/cpp/layout_check.cpp): the
block and fix-up kernels' index structure reproduces a direct scatter exactly
and every index stays within its buffer, for several node orderings."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from xyst_amd import riemann as X
from xyst_amd.meshgen import box_mesh, renumber_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "build", "layout_check")

def _compile(exe, src):
    """g++ into a private name, then an atomic rename: parallel test workers
    never run a half-written binary."""
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    tmp = "%s.%d" % (exe, os.getpid())
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", tmp] + src, check=True)
    os.replace(tmp, exe)


def _exe():
    src = [os.path.join(ROOT, "tests", "cpp", "layout_check.cpp"),
           os.path.join(ROOT, "xyst_amd", "csrc", "xg_layout.cpp")]
    hdr = os.path.join(ROOT, "xyst_amd", "csrc", "xg_layout.hpp")
    if not os.path.exists(EXE) or any(os.path.getmtime(s) > os.path.getmtime(EXE)
                                      for s in src + [hdr]):
        _compile(EXE, src)
    return EXE


def _blocks(kind):
    """(owned nodes, capacity, chunk slots) of a block layout, from the header."""
    import re
    h = open(os.path.join(ROOT, "xyst_amd", "csrc", "xg_layout.hpp")).read()
    return tuple(re.search(r"constexpr int %s = (\d+);" % k, h).group(1)
                 for k in ("k%sNodes" % kind, "k%sCap" % kind, "k%sSlots" % kind))


BLOCKS = {"cho": (), "grad": _blocks("Grad") + ("1",), "rhs": _blocks("Rhs")}   # grad: gather


def _check(tmp_path, m, p, q, blocks=()):
    _, _, besym = X.mesh_symbc(m["coord"], m["triinpoel"], m["faceset"], [1, 2, 3, 4, 5, 6])
    f = tmp_path / "layout.bin"
    with open(f, "wb") as fh:
        np.array([m["coord"].shape[1], p.size, m["triinpoel"].shape[0]], np.int64).tofile(fh)
        p.astype(np.int32).tofile(fh)
        q.astype(np.int32).tofile(fh)
        m["triinpoel"].astype(np.int32).tofile(fh)
        besym.astype(np.uint8).tofile(fh)
    r = subprocess.run([_exe(), str(f), *blocks], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return dict(kv.split("=") for kv in r.stdout.split()[1:])


@pytest.mark.parametrize("blocks", ["cho", "grad", "rhs"])
@pytest.mark.parametrize("order", ["lex", "rcm", "rcm+block"])
def test_layout_matches_direct_scatter(tmp_path, order, blocks):
    m = box_mesh(24, seed=2)
    if order != "lex":
        m = renumber_mesh(m, X.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
    if order == "rcm+block":
        m = renumber_mesh(m, X.mesh_block_order(m["coord"], 2048))
    e, _ = X.mesh_domint(m["coord"], m["inpoel"])
    st = _check(tmp_path, m, e[:, 0].astype(np.int64), e[:, 1].astype(np.int64),
                BLOCKS[blocks])
    if order == "rcm+block":
        assert float(st["halo/owned"]) < {"cho": 0.5, "grad": 2.0, "rhs": 1.1}[blocks]   # grad: two-sided halo


@pytest.mark.parametrize("blocks", ["cho", "grad", "rhs"])
def test_layout_superedge_orientation(tmp_path, blocks):
    """Edges in the mixed orientations of the reference's superedges."""
    m = box_mesh(10, seed=4)
    edge, dint = O.domint(m["coord"], m["inpoel"])
    se = O.domsuped(m["inpoel"], edge, dint)
    s0, s1, s2 = se.sup[0].reshape(-1, 4), se.sup[1].reshape(-1, 3), se.sup[2].reshape(-1, 2)
    lp = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    lt = [(0, 1), (1, 2), (2, 0)]
    P = np.concatenate([s0[:, a] for a, b in lp] + [s1[:, a] for a, b in lt] + [s2[:, 0]])
    Q = np.concatenate([s0[:, b] for a, b in lp] + [s1[:, b] for a, b in lt] + [s2[:, 1]])
    _check(tmp_path, m, P, Q, BLOCKS[blocks])


def test_block_order_is_permutation_of_compact_cells():
    m = box_mesh(16, seed=1)
    perm = X.mesh_block_order(m["coord"], 512)
    assert np.array_equal(np.sort(perm), np.arange(perm.size))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    # runs of 256 consecutive new ids lie within one or two compact cells
    ext = [np.ptp(m["coord"][:, inv[i:i + 256]], axis=1).max()
           for i in range(0, perm.size - 256, 256)]
    assert np.median(ext) < 0.7      # cells of ~512 nodes are about 0.5 wide


TET_EXE = os.path.join(ROOT, "build", "tet_layout_check")


def _tet_exe():
    src = [os.path.join(ROOT, "tests", "cpp", "tet_layout_check.cpp"),
           os.path.join(ROOT, "xyst_amd", "csrc", "xg_layout.cpp")]
    hdr = os.path.join(ROOT, "xyst_amd", "csrc", "xg_layout.hpp")
    if not os.path.exists(TET_EXE) or any(os.path.getmtime(s) > os.path.getmtime(TET_EXE)
                                          for s in src + [hdr]):
        _compile(TET_EXE, src)
    return TET_EXE


@pytest.mark.parametrize("order,tpc", [("lex", 128), ("rcm+block", 128), ("lex", 256),
                                       ("rcm+block", 256)])
def test_tet_layout_matches_direct_scatter(tmp_path, order, tpc):
    """kozak::rhs layout: each tet corner reaches its node exactly once, every
    index the kernel forms is in bounds (CPU emulation, integer fluxes)."""
    m = box_mesh(20, seed=9)
    if order != "lex":
        m = renumber_mesh(m, X.mesh_rcm(m["coord"].shape[1], m["inpoel"]))
        m = renumber_mesh(m, X.mesh_block_order(m["coord"], 2048))
    f = tmp_path / "tets.bin"
    with open(f, "wb") as fh:
        np.array([m["coord"].shape[1], m["inpoel"].shape[0]], np.int64).tofile(fh)
        m["inpoel"].astype(np.int32).tofile(fh)
    r = subprocess.run([_tet_exe(), str(f), str(tpc)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("ok")
