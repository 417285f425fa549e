"""Host side of partitions on one device (CPU only): the partition split of
xyst_amd.partition agrees with the oracle's partitioned restatement, and the
shared-node communicator validates its input before touching the device."""
import ctypes as C

import numpy as np
import pytest

from oracle import partitioned as PP
from tests.analytic_cases import fixture, unitcube
from xyst_amd import _xg
from xyst_amd.partition import split_mesh


def test_split_matches_the_oracle_partitions():
    d = fixture()
    c = unitcube(d)
    owner = d["vf_chare_of_tet"]
    mine = split_mesh(c.coord, c.inpoel, c.tri, owner)
    ref = PP.split(c.coord, c.inpoel, c.tri, owner)
    assert len(mine) == len(ref) == 8
    faces = []
    for a, b in zip(mine, ref):
        assert np.array_equal(a.gid, b.gid)
        assert np.array_equal(a.gid[a.inpoel], c.inpoel[b.tets])
        fa = sorted(tuple(f) for f in a.gid[a.tri])
        fb = sorted(tuple(f) for f in b.tri_glob)
        assert fa == fb
        faces += fa
    # every boundary face belongs to exactly one partition
    assert sorted(faces) == sorted(tuple(f) for f in c.tri)


def test_comm_rejects_bad_input():
    lib = _xg.lib()
    h = C.c_void_p()
    n = np.array([3, 2], np.int64)
    g0 = np.array([0, 1, 2], np.uint64)
    g1 = np.array([2, 2], np.uint64)                  # a repeated global id
    arr = (C.c_void_p * 2)(g0.ctypes.data, g1.ctypes.data)
    assert lib.xg_comm_create(0, 0, n.ctypes.data, arr, C.byref(h)) == _xg.XG_EINVAL
    assert b"partitions" in lib.xg_comm_error(h)
    lib.xg_comm_destroy(h)
    h = C.c_void_p()
    assert lib.xg_comm_create(0, 2, n.ctypes.data, arr, C.byref(h)) == _xg.XG_EINVAL
    assert b"repeats" in lib.xg_comm_error(h)
    lib.xg_comm_destroy(h)
    assert lib.xg_comm_sum_dev(None, 1, None, None) == _xg.XG_EINVAL


def test_partitioned_cg_rejects_bad_input():
    """xg_pcg_create refuses malformed partitions before any device work."""
    from xyst_amd.cg import CSR
    from xyst_amd.partition import PartitionedCG
    inpoel = np.array([[0, 1, 2, 3]], np.int64)
    A = CSR.from_tets(4, inpoel)
    with pytest.raises(_xg.XgError, match="repeats"):
        PartitionedCG([A], [np.array([0, 1, 1, 3])])
    bad = CSR(1, A.ia, A.ja.copy(), A.a)
    bad.ja[-1] = 9
    with pytest.raises(_xg.XgError, match="column"):
        PartitionedCG([bad], [np.arange(4)])
    lib = _xg.lib()
    assert lib.xg_pcg_solve(None, 1, 1e-3, None, None, None) == _xg.XG_EINVAL
    assert lib.xg_pcg_destroy(None) == _xg.XG_EINVAL
