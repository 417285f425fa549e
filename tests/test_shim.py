"""The reference-side C++ shim (xyst_amd/shim: riemann::grad/rhs and
kozak::rhs with the signatures of Riemann.hpp:19-39 and Kozak.hpp:19-27),
compiled against stand-in headers (tests/cpp/shim_stub) into
tests/cpp/libshim_harness.so by the build.

CPU: the shim turns every refused call into a tk::Exception -- a null
context, a host problem with a source term the device was not told about, a
problem whose source the device lacks, mismatched field sizes.  GPU: through
the shim, riemann::grad/rhs and kozak::rhs return exactly what the C-ABI
returns, with and without the vortical-flow source and (kozak::rhs) with
steady local time steps; more scalars than the device serves are refused.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from xyst_amd import _xg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "cpp", "libshim_harness.so")
VP, D, I64 = C.c_void_p, C.c_double, C.c_int64


def _harness():
    _xg.lib()                      # libxg (and torch's HIP runtime) first
    if not os.path.exists(SO):
        pytest.fail("tests/cpp/libshim_harness.so is not built (run the build)")
    h = C.CDLL(SO)
    h.shim_set_src.argtypes = [C.c_int]
    h.shim_set_problem.argtypes = [VP, C.c_char_p, D, D, D, C.c_char_p, C.c_int]
    h.shim_riemann_grad.argtypes = [VP, I64, C.c_int, VP, VP, C.c_char_p, C.c_int]
    h.shim_riemann_rhs.argtypes = [VP, I64, C.c_int, C.c_int, VP, VP, VP, D, VP, VP,
                                   C.c_char_p, C.c_int]
    h.shim_kozak_rhs.argtypes = [VP, I64, C.c_int, D, D, VP, VP, VP, VP, C.c_char_p,
                                 C.c_int]
    return h


def _p(a):
    return None if a is None else a.ctypes.data_as(VP)


class Shim:
    def __init__(self):
        self.h = _harness()
        self.err = C.create_string_buffer(512)

    def _ret(self, rc):
        return None if rc == 0 else self.err.value.decode()

    def set_src(self, on):
        self.h.shim_set_src(1 if on else 0)

    def set_problem(self, ctx, name, alpha=0.1, kappa=1.0, p0=10.0):
        return self._ret(self.h.shim_set_problem(ctx, name.encode(), alpha, kappa, p0,
                                                 self.err, 512))

    def grad(self, ctx, U, nprop=None):
        G = np.zeros((U.shape[0], 3 * U.shape[1]))
        e = self._ret(self.h.shim_riemann_grad(ctx, U.shape[0], U.shape[1], _p(U), _p(G),
                                               self.err, 512))
        return G, e

    def rhs(self, ctx, G, U, v=None, t=0.0, tp=None):
        R = np.zeros_like(U)
        e = self._ret(self.h.shim_riemann_rhs(ctx, U.shape[0], U.shape[1], G.shape[1], _p(G), _p(U),
                                              _p(v), t, _p(tp), _p(R), self.err, 512))
        return R, e

    def kozak(self, ctx, t, dt, U, tp=None, dtp=None):
        R = np.zeros_like(U)
        e = self._ret(self.h.shim_kozak_rhs(ctx, U.shape[0], U.shape[1], t, dt, _p(tp),
                                            _p(dtp), _p(U), _p(R), self.err, 512))
        return R, e


def test_shim_refusals_become_exceptions():
    s = Shim()
    U = np.ones((4, 5))
    G = np.zeros((4, 15))
    s.set_src(False)
    _, e = s.grad(None, U)
    assert e is not None and e.startswith("xg: ")
    _, e = s.rhs(None, G, U)
    assert e is not None and e.startswith("xg: ")
    _, e = s.kozak(None, 0.0, 1e-3, U)
    assert e is not None and e.startswith("xg: ")
    # wrong sizes are caught before the library is called
    _, e = s.rhs(None, np.zeros((4, 10)), U)
    assert "sizes disagree" in e
    # a host problem with a source the device does not evaluate
    s.set_src(True)
    e = s.set_problem(None, "rayleigh_taylor")
    assert e is not None and "not evaluated on the device" in e
    s.set_src(False)


@pytest.mark.gpu
def test_shim_on_the_device_matches_the_c_abi():
    from xyst_amd import kozak as K
    from xyst_amd import riemann as X
    from xyst_amd.meshgen import box_mesh
    s = Shim()
    m = box_mesh(6, seed=4)
    coord, inpoel, tri = m["coord"], m["inpoel"], m["triinpoel"]
    vol = O.vol(coord, inpoel)
    P = O.problem("vortical_flow", gamma=5.0 / 3.0)
    U = O.problem_ic(P, coord)
    U[:, 3] += 0.1 * np.sin(3 * coord[0])
    ctx = X.Context(0, gamma=5.0 / 3.0, flux="hllc")
    edge, dint = X.mesh_domint(coord, inpoel)
    sup, ints = X.edges_as_superedges(edge, dint)
    ctx.upload_mesh(coord, sup, ints, tri, np.zeros(tri.size, np.uint8))
    K.upload(ctx, coord, inpoel)
    h = ctx._h
    s.set_src(False)
    G, e = s.grad(h, U)
    assert e is None
    assert np.array_equal(G, ctx.grad(U))
    Gv = G / vol[:, None]
    R, e = s.rhs(h, Gv, U)
    assert e is None
    assert np.array_equal(R, ctx.rhs(Gv, U))
    # the host problem has a source but the device was not told: refused
    s.set_src(True)
    _, e = s.rhs(h, Gv, U, v=vol)
    assert "none was set on the device" in e
    _, e = s.kozak(h, 0.0, 1e-3, U)
    assert "none was set on the device" in e
    assert s.set_problem(h, "vortical_flow") is None
    R, e = s.rhs(h, Gv, U, v=vol, t=0.5)
    assert e is None
    assert np.array_equal(R, ctx.rhs(Gv, U, v=vol, t=0.5))
    Rk, e = s.kozak(h, 0.0, 1e-3, U)
    assert e is None
    assert np.array_equal(Rk, K.rhs(ctx, 1e-3, U))
    # steady local time steps: each tet takes its corners' mean
    dtp = 1e-3 * (1.0 + 0.5 * np.cos(5 * coord[1]))
    Rs, e = s.kozak(h, 0.0, 1e-3, U, tp=np.zeros(U.shape[0]), dtp=dtp)
    assert e is None
    assert np.array_equal(Rs, K.rhs(ctx, 1e-3, U, dtp=dtp))
    assert not np.array_equal(Rs, Rk)
    # more transported scalars than the device serves: refused
    _, e = s.kozak(h, 0.0, 1e-3, np.ones((U.shape[0], 8)))
    assert e is not None and "scalars" in e
    s.set_src(False)
    ctx.close()
