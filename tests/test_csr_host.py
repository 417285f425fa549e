"""Host side of the pressure-Poisson system through the C-ABI (no GPU):
tk::CSR pattern and ChoCG::prelhs values against the oracle."""
import numpy as np
import pytest

import oracle as O
from tests.test_oracle_cg import _unit_mesh
from xyst_amd import cg as X
from xyst_amd._xg import XgError
from xyst_amd.meshgen import box_mesh


@pytest.mark.parametrize("ncomp", [1, 3])
def test_pattern_matches_oracle(ncomp):
    coord, inpoel, _ = _unit_mesh()
    ia, ja = X.csr_pattern(coord.shape[1], inpoel, ncomp)
    A = O.CSR.from_tets(coord.shape[1], inpoel, ncomp)
    assert np.array_equal(ia.astype(np.int64), A.ia)
    assert np.array_equal(ja.astype(np.int64), A.ja)


def test_pattern_box_mesh():
    m = box_mesh(5, seed=3)
    n = m["coord"].shape[1]
    ia, ja = X.csr_pattern(n, m["inpoel"])
    A = O.CSR.from_tets(n, m["inpoel"])
    assert np.array_equal(ia.astype(np.int64), A.ia)
    assert np.array_equal(ja.astype(np.int64), A.ja)


def test_poisson_matrix_matches_oracle():
    """Same tet order and per-tet arithmetic as the oracle: equal to round-off
    of one tet contribution (the cross products may contract differently)."""
    m = box_mesh(5, seed=3)
    coord, inpoel = m["coord"], m["inpoel"]
    A = X.CSR.poisson(coord, inpoel)
    Ao = O.CSR.from_tets(coord.shape[1], inpoel)
    O.prelhs(coord, inpoel, Ao)
    scale = np.abs(Ao.a).max()
    np.testing.assert_allclose(A.a, Ao.a, rtol=0, atol=1e-14 * scale)
    # and it is -K: symmetric with zero row sums
    M = Ao.dense()
    Mx = np.zeros_like(M)
    for i in range(A.nrow):
        for j in range(A.ia[i], A.ia[i + 1]):
            Mx[i, A.ja[j]] = A.a[j]
    np.testing.assert_allclose(Mx, Mx.T, atol=1e-14 * scale)
    np.testing.assert_allclose(Mx.sum(1), 0.0, atol=1e-13 * scale)


def test_poisson_matrix_rejects_inverted_tet():
    m = box_mesh(2, seed=3)
    inp = m["inpoel"].copy()
    inp[0, [1, 2]] = inp[0, [2, 1]]
    ia, ja = X.csr_pattern(m["coord"].shape[1], inp)
    with pytest.raises(XgError):
        X.poisson_matrix(m["coord"], inp, ia, ja)


def test_pattern_rejects_bad_node_ids():
    inp = np.array([[0, 1, 2, 9]], np.uint64)
    with pytest.raises(XgError):
        X.csr_pattern(4, inp)


def test_chocg_domint_bit_exact_vs_oracle():
    """xg_mesh_chocg_domint (host C++) equals the oracle's ChoCG::domint,
    edge list and all five integrals, bit for bit (same tet-order sums)."""
    import oracle as O
    from xyst_amd import chorin as X
    from xyst_amd.meshgen import box_mesh
    m = box_mesh(7, seed=11)
    e1, d1 = X.mesh_domint(m["coord"], m["inpoel"])
    e0, d0 = O.chocg_domint(m["coord"], m["inpoel"])
    assert np.array_equal(e1.astype(np.int64), e0)
    assert np.array_equal(d1, d0)


@pytest.mark.parametrize("n,seed,stride", [(4, 1, 3), (9, 2, 3), (13, 5, 3), (9, 2, 5)])
def test_domsuped_bit_exact_vs_oracle(n, seed, stride):
    """xg_mesh_domsuped (host C++) groups the edges into the oracle's tet,
    triangle and edge superedges (RieCG::domsuped / ChoCG::domsuped), with
    the same ids, orientations and signed integrals, bit for bit."""
    from xyst_amd import riemann as R
    m = box_mesh(n, seed=seed)
    dom = O.domint if stride == 3 else O.chocg_domint
    edge, dint = dom(m["coord"], m["inpoel"])
    se = O.domsuped(m["inpoel"], edge, dint)
    sup, val = R.mesh_domsuped(m["inpoel"], edge, dint)
    for k in range(3):
        assert np.array_equal(sup[k].astype(np.int64), se.sup[k])
        assert np.array_equal(val[k], se.int[k])
    assert sup[0].size > 0 and sup[1].size > 0
    with pytest.raises(XgError):
        R.mesh_domsuped(m["inpoel"], edge[::-1].copy(), dint)     # unsorted edges
