"""The C-ABI library loads, exports every symbol include/xg.h declares, and
its host-side preprocessing (no GPU needed) matches the oracle bit for bit."""
import os
import re

import numpy as np
import pytest

import oracle as O
from tests.meshes import box_case, sod_fixture
from xyst_amd import _xg
from xyst_amd import riemann as X

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "xg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xg_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _xg.lib()      # loads PyTorch's HIP runtime first, then libxg.so
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_xg.SIG), "binding table out of sync with xg.h"


def test_null_handles_are_rejected():
    """Every entry point that takes a context (or an ExodusII handle) returns
    XG_EINVAL for a null one, with every other argument null or zero, and
    touches nothing: the boundary checks before it dereferences."""
    import ctypes as C
    lib = _xg.lib()
    checked = 0
    for name, (res, args) in _xg.SIG.items():
        if not args or args[0] is not C.c_void_p or res is not C.c_int:
            continue
        if name in ("xg_ctx_destroy",):
            continue
        vals = [0.0 if a is C.c_double else 0 if a in (C.c_int, C.c_int32, C.c_int64)
                else None for a in args]
        assert getattr(lib, name)(*vals) == _xg.XG_EINVAL, name
        checked += 1
    assert checked >= 50
    assert lib.xg_ctx_destroy(None) == _xg.XG_EINVAL
    assert lib.xg_last_error(None)          # a message, not a crash
    assert lib.xg_exo_error(None)
    lib.xg_exo_close(None)
    h = C.c_void_p()
    assert lib.xg_exo_open(None, 0, C.byref(h)) == _xg.XG_EINVAL
    lib.xg_exo_close(h)


def test_context_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_xg.XgError):
        X.Context(0)


@pytest.mark.parametrize("case", ["box", "sod"])
def test_host_preprocessing_matches_oracle(case):
    c = box_case(5, seed=8) if case == "box" else sod_fixture()[0]
    np.testing.assert_array_equal(X.mesh_vol(c.coord, c.inpoel), c.vol)
    e, d = X.mesh_domint(c.coord, c.inpoel)
    np.testing.assert_array_equal(e.astype(np.int64), c.edge)
    np.testing.assert_array_equal(d, c.dint)
    n, nr, b = X.mesh_symbc(c.coord, c.tri, c.faceset, c.symsets)
    np.testing.assert_array_equal(n, c.symnodes)
    np.testing.assert_array_equal(nr, c.symnorms)
    np.testing.assert_array_equal(b, c.besym)


def test_rcm_is_permutation_and_reduces_bandwidth():
    c = box_case(6, seed=2)
    perm = X.mesh_rcm(c.npoin, c.inpoel)
    assert np.array_equal(np.sort(perm), np.arange(c.npoin))
    bw = lambda e: int(np.max(np.abs(e[:, 0] - e[:, 1])))
    # a random relabelling has a large bandwidth; RCM restores a small one
    rnd = np.random.default_rng(0).permutation(c.npoin)
    inp = rnd[c.inpoel]
    e0 = rnd[c.edge]
    p2 = X.mesh_rcm(c.npoin, inp)
    assert bw(p2[e0]) < bw(e0) / 4


def test_error_codes():
    with pytest.raises(_xg.XgError):
        X.mesh_vol(np.zeros((3, 4)), np.array([[0, 1, 2, 3]]))     # J = 0
    c = box_case(2)
    bad = c.inpoel.copy()
    bad[0, 0] = c.npoin + 5
    with pytest.raises(_xg.XgError):
        X.mesh_domint(c.coord, bad)
