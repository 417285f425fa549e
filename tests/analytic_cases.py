"""The reference's analytic-problem regression cases (RieCG/VorticalFlow,
RieCG/TaylorGreen, RieCG/Pipe, KozCG/VorticalFlow, KozCG/TaylorGreen) as
inputs and comparison rules shared by the oracle tests and the GPU tests.

Fixture: tests/golden/riecg_analytic.npz (make_regression_fixtures.py).
Diagnostics rows follow NodeDiagnostics/Transporter (NodeDiagnostics.cpp:
46-146, Transporter.cpp:1472-1494): it, t, dt, L2(U_c), L2(U_c - Un_c), the
total energy, then with an analytic solution L2 and L1 of the primitive
error.
"""
import os

import numpy as np

import oracle as O
from tests.meshes import Case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GAMMA_VF = 5.0 / 3.0


def fixture():
    return np.load(os.path.join(GOLDEN, "riecg_analytic.npz"))


def unitcube(d):
    return Case(d["coord"], d["inpoel"], d["triinpoel"], d["faceset"], [])


def pipe_case(d):
    return Case(d["pipe_coord"], d["pipe_inpoel"], d["pipe_triinpoel"],
                d["pipe_faceset"], [2, 4, 5, 6])


def all_dirichlet(c, ncomp=5):
    """bc_dir = {{1,1,1,1,1,1} ... {6,...}}: every component of every boundary
    node set to the initial condition."""
    return O.dirbc_lists(c.tri, c.faceset, ncomp, [(s, [1] * ncomp) for s in range(1, 7)])


def diag_row(it, t, dt, dg, V, de=None):
    row = [it, t, dt] + list(np.sqrt(dg[:10] / V)) + [dg[10]]
    if de is not None:
        row += list(np.sqrt(de[:5] / V)) + list(de[5:] / V)
    return np.array(row)


def primitive_fields(U, g):
    """density, velocity, specific total energy, pressure (the fields the
    reference writes)."""
    r = U[:, 0]
    u, v, w, e = U[:, 1] / r, U[:, 2] / r, U[:, 3] / r, U[:, 4] / r
    p = (g - 1.0) * r * (e - 0.5 * (u * u + v * v + w * w))
    return np.stack([r, u, v, w, e, p], 1)


def check_diag(mine, ref, tight=True):
    """diag.ndiff.cfg of RieCG/VorticalFlow (cols 2-8 any abs 3e-5 rel 2e-5,
    cols 9-13 any abs 1e-2 rel 1e-7, 14- ignored) and, tighter, agreement
    with the printed 9 digits: conserved L2 norms, t, dt and the total
    energy to 2e-8, the analytic-error norms to 5e-7, the residual norms
    (differences of nearby states) to 1e-5."""
    rel = np.abs(mine - ref) / np.maximum(np.abs(ref), 1e-300)
    ab = np.abs(mine - ref)
    ok = (ab[1:8] <= 3e-5) | (rel[1:8] <= 2e-5)
    assert ok.all(), ("cols 2-8", rel[1:8])
    ok = (ab[8:13] <= 1e-2) | (rel[8:13] <= 1e-7)
    assert ok.all(), ("cols 9-13", rel[8:13])
    if tight:
        assert (rel[1:8] <= 2e-8).all(), ("cols 2-8 tight", rel[1:8])
        assert rel[13] <= 2e-8, ("total energy", rel[13])
        assert (rel[8:13] <= 1e-5).all(), ("residuals", rel[8:13])
        if ref.size > 14:
            assert (rel[14:] <= 5e-7).all(), ("error norms", rel[14:])


def check_fields(mine, ref, bands):
    """exodiff-style comparison: per field (relative tol, floor) or
    ('abs', tol, floor)."""
    for i, (kind, tol, floor) in enumerate(bands):
        a, b = mine[:, i], ref[:, i]
        if kind == "rel":
            ok = (np.abs(a - b) <= tol * np.abs(b)) | (np.abs(b) < floor)
        else:
            ok = (np.abs(a - b) <= tol) | (np.abs(b) < floor)
        assert ok.all(), (i, np.abs(a - b).max())


# exodiff.cfg of RieCG/VorticalFlow (KozCG/VorticalFlow: velocityz 7e-6)
VF_BANDS = [("rel", 1e-7, 1e-8), ("rel", 7e-6, 1e-8), ("rel", 4e-6, 1e-8),
            ("rel", 2e-5, 1e-8), ("rel", 1e-7, 1e-8), ("rel", 1e-7, 1e-8)]
# exodiff.cfg of RieCG/TaylorGreen and KozCG/TaylorGreen
TG_BANDS = [("rel", 1e-7, 1e-8), ("rel", 2e-7, 1e-8), ("rel", 4e-7, 1e-8),
            ("rel", 3e-5, 1e-8), ("rel", 1e-7, 1e-8), ("rel", 1e-7, 1e-8)]
# exodiff.cfg of RieCG/Pipe
PIPE_BANDS = [("rel", 1e-6, 1e-7), ("abs", 1e-6, 0.0), ("abs", 2e-7, 1e-6),
              ("abs", 2e-7, 1e-6), ("rel", 1e-6, 1e-7), ("rel", 1e-6, 1e-7)]


# ---- RieCG/SlotCyl: one transported scalar -----------------------------------

def slotcyl_fixture():
    return np.load(os.path.join(GOLDEN, "riecg_slotcyl.npz"))


def slotcyl_case(d):
    return Case(d["coord"], d["inpoel"], d["triinpoel"], d["faceset"], [])


def slotcyl_dirichlet(c):
    """bc_dir of slot_cyl.q: side set 1 all six components, side set 2 the
    five flow components (c0 left free)."""
    return O.dirbc_lists(c.tri, c.faceset, 6, [(1, [1] * 6), (2, [1] * 5 + [0])])


def slotcyl_diag_row(it, t, dt, dg, V, de):
    """it, t, dt, L2(U_c) (6), L2(U_c - Un_c) (6), total energy, L2 and L1
    of the error (6 each)."""
    return np.array([it, t, dt] + list(np.sqrt(dg[:12] / V)) + [dg[12]]
                    + list(np.sqrt(de[:6] / V)) + list(de[6:] / V))


def check_slotcyl_diag(mine, ref, tight):
    """diag.ndiff.cfg of RieCG/SlotCyl: cols 2-9 any abs 1e-4 rel 1e-7,
    cols 10-15 any abs 6e-5 rel 1e-7, col 16 rel 1e-7, the error columns
    ignored.  tight: agreement with the printed 12 digits (t, dt, the
    conserved and residual norms and the total energy to 1e-9; the error
    norms to 1e-7)."""
    rel = np.abs(mine - ref) / np.maximum(np.abs(ref), 1e-300)
    ab = np.abs(mine - ref)
    assert ((ab[1:9] <= 1e-4) | (rel[1:9] <= 1e-7)).all(), ("cols 2-9", ab[1:9])
    assert ((ab[9:15] <= 6e-5) | (rel[9:15] <= 1e-7)).all(), ("cols 10-15", ab[9:15])
    assert rel[15] <= 1e-7, ("total energy", rel[15])
    if tight:
        nz = np.abs(ref) > 0
        assert (rel[1:16][nz[1:16]] <= 1e-9).all(), ("tight", rel[1:16])
        assert (ab[1:16][~nz[1:16]] <= 1e-15).all(), ("tight zeros", ab[1:16])
        assert (rel[16:][nz[16:]] <= 1e-7).all(), ("errors", rel[16:])


def slotcyl_fields(U, g):
    """The seven fields of slot_cyl.std.exo: primitive flow fields and c0."""
    return np.concatenate([primitive_fields(U[:, :5], g), U[:, 5:6]], 1)


# exodiff.cfg of RieCG/SlotCyl: relative 1e-7 floor 1e-8, zvelocity absolute
# 1e-7 floor 1e-7, c0 absolute 1e-8
SC_BANDS = [("rel", 1e-7, 1e-8), ("rel", 1e-7, 1e-8), ("rel", 1e-7, 1e-8),
            ("abs", 1e-7, 1e-7), ("rel", 1e-7, 1e-8), ("rel", 1e-7, 1e-8),
            ("abs", 1e-8, 0.0)]
