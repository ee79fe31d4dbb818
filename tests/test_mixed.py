"""Distinct trial and test bases on the CPU: the host's argument checks
(the reference's quadrature-count error, the combinations that are not
accelerated), and the oracle's mixed assembly pinned by the mathematics --
swapping the trial and test bases of a symmetric form transposes the
matrix, and since P1 is a subspace of P2 (a P1 field is the P2 field with
its vertex values and the edge means at the midpoints), the P2-trial /
P1-test matrix applied to the embedded field equals the P1 matrix applied
to the P1 field."""
import numpy as np
import pytest

import femhip
from femhip.models import laplace, mass
from oracle import skfem_np as O


def _tri(n=6, a=0.2):
    x = np.linspace(0, 1, n)
    m = O.MeshTri.init_tensor(x, x)
    return O.jitter(m, a, 1.0 / (n - 1)) if a else m


def _tet(n=4, a=0.2):
    x = np.linspace(0, 1, n)
    m = O.MeshTet.init_tensor(x, x, x)
    return O.jitter(m, a, 1.0 / (n - 1)) if a else m


def embed_p1_in_p2(b1, b2, w1):
    """The P2 DOF vector of the P1 field w1: vertex DOFs copied, every edge
    DOF the mean of its edge's vertex values (the local P2 order lists the
    vertices, then the edges of the reference element)."""
    edges = O.TRI_FACETS if b2.mesh.kind == 'tri' else O.TET_EDGES
    nv = b1.Nbfun
    w2 = np.zeros(b2.N)
    ed1, ed2 = b1.element_dofs, b2.element_dofs
    for k in range(nv):
        w2[ed2[k]] = w1[ed1[k]]
    for k, (a, b) in enumerate(edges):
        w2[ed2[nv + k]] = 0.5 * (w1[ed1[a]] + w1[ed1[b]])
    return w2


@pytest.mark.parametrize('mk,E1,E2', [(_tri, 'ElementTriP1', 'ElementTriP2'),
                                      (_tet, 'ElementTetP1', 'ElementTetP2')])
@pytest.mark.parametrize('form', ['laplace', 'mass'])
def test_oracle_mixed_assembly_is_pinned(mk, E1, E2, form):
    m = mk()
    f = {'laplace': O.form_laplace, 'mass': O.form_mass}[form]
    b1 = O.CellBasis(m, getattr(O, E1)(), intorder=4)
    b2 = O.CellBasis(m, getattr(O, E2)(), intorder=4)
    A12 = O.bilinear_assemble(f, b2, vbasis=b1)     # trial P2, test P1
    A21 = O.bilinear_assemble(f, b1, vbasis=b2)     # trial P1, test P2
    assert A12.shape == (b1.N, b2.N) and A21.shape == (b2.N, b1.N)
    T = A21.T.tocsr()
    T.sort_indices()
    assert np.array_equal(T.indptr, A12.indptr) and np.array_equal(T.indices, A12.indices)
    assert np.abs(T.data - A12.data).max() <= 1e-14 * np.abs(A12.data).max()
    A11 = O.bilinear_assemble(f, b1)
    w1 = np.random.default_rng(0).standard_normal(b1.N)
    w2 = embed_p1_in_p2(b1, b2, w1)
    r = A11 @ w1
    assert np.abs(A12 @ w2 - r).max() <= 1e-12 * np.abs(r).max()
    # the one-basis call is the mixed call with the same basis twice
    B = O.bilinear_assemble(f, b1, vbasis=b1)
    assert np.array_equal(B.indices, A11.indices) and np.array_equal(B.data, A11.data)


def test_oracle_quadrature_mismatch():
    m = _tri()
    with pytest.raises(ValueError, match="Quadrature mismatch"):
        O.bilinear_coo(O.form_mass, O.CellBasis(m, O.ElementTriP1()),
                       vbasis=O.CellBasis(m, O.ElementTriP2()))


def test_mixed_argument_checks():
    x = np.linspace(0, 1, 5)
    m = femhip.MeshTri.init_tensor(x, x)
    b1 = femhip.CellBasis(m, femhip.ElementTriP1())
    b2 = femhip.CellBasis(m, femhip.ElementTriP2())
    with pytest.raises(ValueError, match="Quadrature mismatch"):
        mass.assemble(b1, b2)                    # default orders 2 and 4
    b1q = femhip.CellBasis(m, femhip.ElementTriP1(), intorder=4)
    vb = femhip.CellBasis(m, femhip.ElementVector(femhip.ElementTriP2()))
    with pytest.raises(NotImplementedError):
        mass.assemble(vb, b1q)                   # vector trial, scalar test
    sub = femhip.CellBasis(m, femhip.ElementTriP2(), elements=np.arange(5))
    with pytest.raises(NotImplementedError):
        laplace.assemble(b2, sub)                # different elements
    b2o = femhip.CellBasis(m, femhip.ElementTriP2(), intorder=3)
    b1o = femhip.CellBasis(m, femhip.ElementTriP1(), intorder=4)
    if b2o.X.shape[-1] == b1o.X.shape[-1]:
        with pytest.raises(NotImplementedError):
            laplace.assemble(b2o, b1o)           # same count, other points
    with pytest.raises(ValueError):
        femhip.asm(femhip.models.unit_load, b1, b1)
