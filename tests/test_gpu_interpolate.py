"""Basis.interpolate on the device: values and gradients of a DOF vector at
the quadrature points, bit-identical to the oracle's restatement of the
reference, and its use as a coefficient (a 1-D DOF array passed as
w['kappa'] is interpolated, as the reference does)."""
import numpy as np
import pytest

import femhip
from femhip.models import kappa_laplace
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu


def _mesh(kind, n, a):
    x = np.linspace(0, 1, n)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet, 'hex': femhip.MeshHex}[kind]
    m = M.init_tensor(x, x) if kind == 'tri' else M.init_tensor(x, x, x)
    return m.jittered(a, 1.0 / (n - 1)) if a else m


def _omesh(m):
    return {'tri': O.MeshTri, 'tet': O.MeshTet, 'hex': O.MeshHex}[m.kind](m.p, m.t)


@pytest.mark.parametrize('kind,elem,n,a', [
    ('tri', 'ElementTriP1', 17, 0.2), ('tri', 'ElementTriP2', 9, 0.2),
    ('tet', 'ElementTetP1', 7, 0.2), ('tet', 'ElementTetP2', 5, 0.2),
    ('hex', 'ElementHex1', 5, 0.15), ('hex', 'ElementHex2', 4, 0.15),
])
def test_interpolate_bitexact(kind, elem, n, a):
    m = _mesh(kind, n, a)
    b = femhip.Basis(m, getattr(femhip, elem)())
    w = np.random.default_rng(7).standard_normal(b.N)
    print('case', kind, elem, n, flush=True)
    f = b.interpolate(w)
    ob = O.CellBasis(_omesh(m), getattr(O, elem)())
    val, grad = O.interpolate(ob, w)
    assert f.value.shape == val.shape and f.grad.shape == grad.shape
    assert np.array_equal(f.value, val)
    assert np.array_equal(f.grad, grad)


@pytest.mark.parametrize('kind,elem,n', [('tri', 'ElementTriP1', 17),
                                         ('tet', 'ElementTetP2', 5)])
def test_dof_vector_coefficient(kind, elem, n):
    """kappa given as a DOF vector of the same basis: the reference
    interpolates it (form.py:106-108); pattern bit-exact, values 1e-12."""
    m = _mesh(kind, n, 0.2)
    b = femhip.Basis(m, getattr(femhip, elem)())
    w = 1.5 + np.sin(3 * np.arange(b.N))
    A = kappa_laplace.assemble(b, kappa=w)
    ob = O.CellBasis(_omesh(m), getattr(O, elem)())
    kq, _ = O.interpolate(ob, w)
    R = O.assemble_named('kappa_laplace', ob, kappa=kq)
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert np.abs(A.data - R.data).max() <= 1e-12 * np.abs(R.data).max()


def test_per_point_coefficient():
    """An (E, Q) coefficient whose columns differ enters per quadrature
    point (regression: it must not be read as per-element)."""
    m = _mesh('tri', 17, 0.2)
    b = femhip.Basis(m, femhip.ElementTriP2())
    kq = np.random.default_rng(9).lognormal(0, 1, (m.nelements, b.W.shape[0]))
    A = kappa_laplace.assemble(b, kappa=kq)
    ob = O.CellBasis(_omesh(m), O.ElementTriP2())
    R = O.assemble_named('kappa_laplace', ob, kappa=kq)
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert np.abs(A.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
