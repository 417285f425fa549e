"""CPU checks of the AMG setup (tests/cpp/amg_check.cpp over
xyst_amd/csrc/xg_amg.cpp): Galerkin coarse operators, their symmetry, and
multigrid-preconditioned CG on the ChoCG pressure matrix (one Dirichlet face,
or a single hydrostat node) converging in a mesh-independent number of
iterations."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from xyst_amd.meshgen import box_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "build", "amg_check")


def _exe():
    src = [os.path.join(ROOT, "tests", "cpp", "amg_check.cpp"),
           os.path.join(ROOT, "xyst_amd", "csrc", "xg_amg.cpp")]
    hdr = [os.path.join(ROOT, "xyst_amd", "csrc", "xg_amg.hpp")]
    if not os.path.exists(EXE) or any(os.path.getmtime(s) > os.path.getmtime(EXE)
                                      for s in src + hdr):
        os.makedirs(os.path.dirname(EXE), exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", EXE] + src, check=True)
    return EXE


def _system(n, dirichlet):
    m = box_mesh(n, seed=6)
    coord, inpoel = m["coord"], m["inpoel"]
    A = O.CSR.from_tets(coord.shape[1], inpoel)
    O.prelhs(coord, inpoel, A)
    nodes = np.where(coord[0] < 1e-12)[0] if dirichlet == "face" else np.array([0])
    rr = np.zeros(coord.shape[1])
    for i in sorted(nodes, reverse=True):
        A.dirichlet(int(i), 0.0, rr)
    b = np.sin(3 * coord[0]) * np.cos(2 * coord[1]) + coord[2]
    b[nodes] = 0.0
    return A, b


def _run(tmp_path, A, b):
    f = tmp_path / "amg.bin"
    n = A.ia.size - 1
    with open(f, "wb") as fh:
        np.array([n, A.ja.size], np.int64).tofile(fh)
        A.ia.astype(np.int64).tofile(fh)
        A.ja.astype(np.int32).tofile(fh)
        A.a.astype(np.float64).tofile(fh)
        b.astype(np.float64).tofile(fh)
    r = subprocess.run([_exe(), str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return dict(kv.split("=") for kv in r.stdout.split()[1:] if "=" in kv)


@pytest.mark.parametrize("n,dirichlet", [(8, "face"), (16, "face"), (16, "hydrostat")])
def test_amg_setup_and_preconditioned_cg(tmp_path, n, dirichlet):
    A, b = _system(n, dirichlet)
    st = _run(tmp_path, A, b)
    assert float(st["relres"]) <= 1e-10
    assert float(st["complexity"]) < 1.8
    assert int(st["iterations"]) <= 25, st
