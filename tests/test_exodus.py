"""ExodusII reader/writer (xg_exo_*, xyst_amd.exodus) on the host.

Pinned three ways: against scipy.io.netcdf_file, an independent NetCDF
classic implementation, on the reference's own files (tests/golden/exo,
copied unchanged from its regression tests); against the RieCG/Sod fixture
mesh, which was read with scipy and oriented by the partitioner rule; and by
round trips of files this writer makes, read back by both readers.
"""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from xyst_amd import exodus
from xyst_amd._xg import lib
from xyst_amd.meshgen import box_mesh

HERE = os.path.dirname(os.path.abspath(__file__))
EXO = os.path.join(HERE, "golden", "exo")
FILES = ["rectangle_01_1.5k.exo", "unitcube_1k.exo", "ldc.std.exo"]


def scipy_read(path):
    return netcdf_file(path, "r", mmap=False)


@pytest.mark.parametrize("name", FILES)
def test_reference_files_match_scipy(name):
    path = os.path.join(EXO, name)
    f = scipy_read(path)
    m = exodus.read_mesh(path, orient=False)
    X = np.stack([f.variables[c].data for c in ("coordx", "coordy", "coordz")])
    assert np.array_equal(m["coord"], X)
    # tetrahedra: every 4-node block, in block order
    tets, tris = [], []
    k = 1
    while "connect%d" % k in f.variables:
        c = f.variables["connect%d" % k].data.astype(np.int64) - 1
        (tets if c.shape[1] == 4 else tris).append(c)
        k += 1
    assert np.array_equal(m["inpoel"], np.concatenate(tets))
    # side sets resolved to triangles (triangle elements or tetrahedron sides)
    nss = f.dimensions.get("num_side_sets", 0) or 0
    assert np.unique(m["faceset"]).size == nss
    if nss:
        ids = f.variables["ss_prop1"].data
        tri_blk = tris[0] if tris else None
        tet_blk = np.concatenate(tets)
        side_nodes = {1: (0, 1, 3), 2: (1, 2, 3), 3: (0, 3, 2), 4: (0, 2, 1)}
        want, wset = [], []
        for s in range(1, nss + 1):
            elem = f.variables["elem_ss%d" % s].data.astype(np.int64)
            side = f.variables["side_ss%d" % s].data.astype(np.int64)
            for e, sd in zip(elem, side):
                if tri_blk is not None and e <= tri_blk.shape[0]:
                    want.append(tri_blk[e - 1])
                else:
                    e0 = e - 1 - (tri_blk.shape[0] if tri_blk is not None else 0)
                    want.append(tet_blk[e0][list(side_nodes[int(sd)])])
            wset += [ids[s - 1]] * elem.size
        assert np.array_equal(m["triinpoel"], np.array(want))
        assert np.array_equal(m["faceset"], np.array(wset))


def test_reference_result_fields_match_scipy():
    path = os.path.join(EXO, "ldc.std.exo")
    f = scipy_read(path)
    names = [b"".join(x).decode().strip("\x00 ") for x in f.variables["name_nod_var"].data]
    t = f.variables["time_whole"].data
    assert np.array_equal(exodus.times(path), t)
    for step in range(t.size):
        nm, fields, time = exodus.read_fields(path, step)
        assert nm == names and time == t[step]
        for v in range(len(names)):
            assert np.array_equal(fields[:, v], f.variables["vals_nod_var%d" % (v + 1)].data[step])
    nm, fields, time = exodus.read_fields(path)      # last step
    assert time == t[-1]


def test_sod_mesh_matches_fixture():
    """The RieCG/Sod mesh as the regression fixture holds it (read there with
    scipy, triangles oriented by the partitioner rule)."""
    m = exodus.read_mesh(os.path.join(EXO, "rectangle_01_1.5k.exo"))
    F = np.load(os.path.join(HERE, "golden", "riecg_sod.npz"))
    assert np.array_equal(m["coord"], F["coord"])
    assert np.array_equal(m["inpoel"], F["inpoel"])
    assert np.array_equal(m["triinpoel"], F["triinpoel"])
    assert np.array_equal(m["faceset"], F["faceset"])


def _write_box(path, n=3, nstep=3):
    m = box_mesh(n, seed=7)
    N = m["coord"].shape[1]
    rng = np.random.default_rng(1)
    steps = [(0.1 * (k + 1), rng.standard_normal((N, 2))) for k in range(nstep)]
    with exodus.Writer(path, m, names=["density", "pressure"]) as w:
        for t, f in steps:
            w.write_step(t, f)
    return m, steps


def test_writer_round_trip_and_scipy_reads_it(tmp_path):
    path = str(tmp_path / "box.exo")
    m, steps = _write_box(path)
    r = exodus.read_mesh(path, orient=False)
    assert np.array_equal(r["coord"], m["coord"])
    assert np.array_equal(r["inpoel"], m["inpoel"])
    # side sets come back grouped by id, each in the input order
    order = np.argsort(m["faceset"], kind="stable")
    assert np.array_equal(r["triinpoel"], m["triinpoel"][order])
    assert np.array_equal(r["faceset"], m["faceset"][order])
    for k, (t, f) in enumerate(steps):
        names, fields, time = exodus.read_fields(path, k)
        assert names == ["density", "pressure"] and time == t
        assert np.array_equal(fields, f)
    # an independent NetCDF reader sees the same file
    g = scipy_read(path)
    assert g.version_byte == 2
    assert np.array_equal(g.variables["coordy"].data, m["coord"][1])
    assert np.array_equal(g.variables["connect2"].data - 1, m["inpoel"])
    assert g.variables["connect2"].elem_type == b"TETRA4"
    assert np.array_equal(g.variables["time_whole"].data, [t for t, _ in steps])
    assert np.array_equal(g.variables["vals_nod_var2"].data[1], steps[1][1][:, 1])
    assert list(g.variables["ss_prop1"].data) == sorted(set(m["faceset"].tolist()))


def test_append_mode_adds_steps(tmp_path):
    path = str(tmp_path / "box.exo")
    m, steps = _write_box(path, nstep=2)
    N = m["coord"].shape[1]
    f = np.arange(2 * N, dtype=np.float64).reshape(N, 2)
    with exodus.Writer(path, append=True) as w:
        assert w.nstep == 2
        w.write_step(5.0, f)
    assert np.array_equal(exodus.times(path), [0.1, 0.2, 5.0])
    names, fields, time = exodus.read_fields(path)
    assert time == 5.0 and np.array_equal(fields, f)
    assert np.array_equal(exodus.read_fields(path, 0)[1], steps[0][1])
    g = scipy_read(path)
    assert g.variables["vals_nod_var1"].data.shape == (3, N)


def test_cdf1_file_from_scipy(tmp_path):
    """A CDF-1 (32-bit offset) file written by scipy, with sides of
    tetrahedra and a single record variable (unpadded record size)."""
    path = str(tmp_path / "cdf1.exo")
    m = box_mesh(2, seed=3)
    N, T = m["coord"].shape[1], m["inpoel"].shape[0]
    g = netcdf_file(path, "w", version=1)
    g.createDimension("time_step", None)
    g.createDimension("num_nodes", N)
    g.createDimension("num_el_blk", 1)
    g.createDimension("num_el_in_blk1", T)
    g.createDimension("num_nod_per_el1", 4)
    g.createDimension("num_side_sets", 1)
    g.createDimension("num_side_ss1", 2)
    g.createDimension("num_nod_var", 1)
    g.createDimension("len_name", 33)
    for j, c in enumerate("xyz"):
        v = g.createVariable("coord" + c, "d", ("num_nodes",))
        v[:] = m["coord"][j]
    v = g.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
    v.elem_type = b"TETRA"
    v[:] = m["inpoel"] + 1
    g.createVariable("ss_prop1", "i", ("num_side_sets",))[:] = [7]
    g.createVariable("elem_ss1", "i", ("num_side_ss1",))[:] = [1, 2]
    g.createVariable("side_ss1", "i", ("num_side_ss1",))[:] = [4, 1]
    nm = g.createVariable("name_nod_var", "c", ("num_nod_var", "len_name"))
    nm[0, :3] = np.array(list("rho"), "S1")
    vv = g.createVariable("vals_nod_var1", "d", ("time_step", "num_nodes"))
    vv[0] = np.arange(N) * 0.5
    vv[1] = np.arange(N) * -1.0
    g.close()
    r = exodus.read_mesh(path, orient=False)
    assert np.array_equal(r["coord"], m["coord"])
    assert np.array_equal(r["inpoel"], m["inpoel"])
    t0, t1 = m["inpoel"][0], m["inpoel"][1]
    assert np.array_equal(r["triinpoel"], [t0[[0, 2, 1]], t1[[0, 1, 3]]])
    assert list(r["faceset"]) == [7, 7]
    col = np.empty(N)
    with exodus._Handle(path) as h:      # no time_whole: read the variable directly
        assert h.sizes()["nstep"] == 2
        h.check(lib().xg_exo_read_node_var(h.h, 1, 0, col.ctypes.data))
    assert np.array_equal(col, np.arange(N) * -1.0)


def test_errors(tmp_path):
    with pytest.raises(exodus.ExodusError, match="cannot open"):
        exodus.read_mesh(str(tmp_path / "missing.exo"))
    bad = tmp_path / "bad.exo"
    bad.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(exodus.ExodusError, match="not a NetCDF classic"):
        exodus.read_mesh(str(bad))
    good = open(os.path.join(EXO, "unitcube_1k.exo"), "rb").read()
    trunc = tmp_path / "trunc.exo"
    trunc.write_bytes(good[:200])
    with pytest.raises(exodus.ExodusError, match="malformed or truncated"):
        exodus.read_mesh(str(trunc))
    cut = tmp_path / "cut.exo"
    cut.write_bytes(good[: len(good) // 2])
    with pytest.raises(exodus.ExodusError, match="past the end"):
        exodus.read_mesh(str(cut))
    path = str(tmp_path / "box.exo")
    _write_box(path, nstep=1)
    with pytest.raises(exodus.ExodusError, match="step out of range"):
        exodus.read_fields(path, 3)
    m = box_mesh(2, seed=1)
    m["inpoel"] = m["inpoel"].copy()
    m["inpoel"][0, 0] = 10 ** 6
    with pytest.raises(exodus.ExodusError, match="node id out of range"):
        exodus.Writer(str(tmp_path / "x.exo"), m, names=["a"])
    with exodus.Writer(path, append=True) as w:
        with pytest.raises(exodus.ExodusError, match="expected 2 node variables"):
            w.write_step(1.0, np.zeros((w.npoin, 3)))


def test_connect_size_mismatch_is_an_error(tmp_path):
    """A connect variable whose size disagrees with its block's
    num_el_in_blk * num_nod_per_el (a malformed file) is rejected instead of
    overflowing the caller's buffer, which is sized from the block."""
    path = str(tmp_path / "mismatch.exo")
    m = box_mesh(2, seed=3)
    N, T = m["coord"].shape[1], m["inpoel"].shape[0]
    g = netcdf_file(path, "w", version=1)
    g.createDimension("num_nodes", N)
    g.createDimension("num_el_blk", 1)
    g.createDimension("num_el_in_blk1", T - 4)        # block says T - 4 tets
    g.createDimension("num_nod_per_el1", 4)
    g.createDimension("num_tets_stored", T)            # connect holds T
    for j, c in enumerate("xyz"):
        v = g.createVariable("coord" + c, "d", ("num_nodes",))
        v[:] = m["coord"][j]
    v = g.createVariable("connect1", "i", ("num_tets_stored", "num_nod_per_el1"))
    v.elem_type = b"TETRA"
    v[:] = m["inpoel"] + 1
    g.close()
    with pytest.raises(exodus.ExodusError, match="disagrees with its block"):
        exodus.read_mesh(path, orient=False)
