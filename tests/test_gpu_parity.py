"""GPU parity of the assembly path against the reference's golden fixtures
and the oracle (pattern bit-exact, values within 1e-12 of max|A|)."""
import glob
import json
import os

import numpy as np
import pytest

import femhip
from femhip.models import (laplace, mass, unit_load, kappa_laplace,
                           linear_elasticity, lame_parameters, vector_laplace)
from oracle import skfem_np as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
TOL = 1e-12

ELEM = {'tri1': femhip.ElementTriP1, 'tri2': femhip.ElementTriP2,
        'tet1': femhip.ElementTetP1, 'tet2': femhip.ElementTetP2,
        'hex1': femhip.ElementHex1, 'hex2': femhip.ElementHex2}
MESH = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet, 'hex': femhip.MeshHex}

pytestmark = pytest.mark.gpu


def _fixtures():
    return sorted(glob.glob(os.path.join(GOLD, 'small_*.npz')))


def _basis(z, meta):
    m = MESH[meta['kind']](z['p'], z['t'])
    e = ELEM[meta['element']]()
    if meta['vector']:
        e = femhip.ElementVector(e)
    io = None if meta['intorder'] < 0 else meta['intorder']
    return femhip.Basis(m, e, intorder=io)


def _assemble(form, b, z):
    if form == 'laplace':
        return laplace.assemble(b)
    if form == 'mass':
        return mass.assemble(b)
    if form == 'unit_load':
        return unit_load.assemble(b)
    if form == 'kappa_laplace':
        return kappa_laplace.assemble(b, kappa=z['kappa_e'][:, None])
    if form == 'linear_elasticity':
        lam, mu = lame_parameters(1.0, 0.3)
        return linear_elasticity(lam, mu).assemble(b)
    if form == 'vector_laplace':
        return vector_laplace.assemble(b)
    raise KeyError(form)


@pytest.mark.parametrize('path', _fixtures(), ids=lambda p: os.path.basename(p)[6:-4])
def test_small_fixture(path):
    z = np.load(path)
    meta = json.loads(str(z['meta']))
    name = os.path.basename(path)[6:-4]
    b = _basis(z, meta)
    assert np.array_equal(b.element_dofs, z['element_dofs'])
    print('case', name, flush=True)
    for form in meta['forms']:
        r = _assemble(form, b, z)
        if form == 'unit_load':
            ref = z['unit_load_b']
            assert r.shape == ref.shape
            assert np.abs(r - ref).max() <= TOL * np.abs(ref).max()
            continue
        assert r.indptr.dtype == np.int32 and r.indices.dtype == np.int32
        assert r.has_sorted_indices
        assert np.array_equal(r.indptr, z[form + '_indptr']), (name, form)
        assert np.array_equal(r.indices, z[form + '_indices']), (name, form)
        ref = z[form + '_data']
        err = np.abs(r.data - ref).max() / np.abs(ref).max()
        assert err <= TOL, (name, form, err)


def test_config1_tri129_with_enforce():
    z = np.load(os.path.join(GOLD, 'config1_tri129.npz'))
    x = np.linspace(0, 1, 129)
    m = femhip.MeshTri.init_tensor(x, x)
    b = femhip.Basis(m, femhip.ElementTriP1())
    A = laplace.assemble(b)
    f = unit_load.assemble(b)
    assert A.nnz == 82689
    assert np.array_equal(A.indptr, z['indptr'])
    assert np.array_equal(A.indices, z['indices'])
    assert np.abs(A.data - z['data']).max() <= TOL * np.abs(z['data']).max()
    assert np.abs(f - z['b']).max() <= TOL * np.abs(z['b']).max()
    D = m.boundary_nodes()
    assert np.array_equal(D, z['D'])
    A2, f2 = femhip.enforce(A, f, D=D)
    Ar, fr = O.enforce(_csr(z), z['b'], D=z['D'])
    assert np.array_equal(A2.indices, Ar.indices)
    assert np.abs(A2.data - Ar.data).max() <= TOL * np.abs(Ar.data).max()
    assert np.array_equal(f2, fr) or np.abs(f2 - fr).max() <= TOL


def _csr(z):
    from scipy.sparse import csr_matrix
    n = len(z['indptr']) - 1
    return csr_matrix((z['data'], z['indices'], z['indptr']), shape=(n, n))


def test_ex01_max():
    """docs/examples/ex01.py: P1 Laplace + load + enforce on the 64x64
    refined unit square; tests/test_examples.py:8-12 pins max(x)."""
    from scipy.sparse.linalg import spsolve
    x = np.linspace(0, 1, 65)
    m = femhip.MeshTri.init_tensor(x, x)
    b = femhip.Basis(m, femhip.ElementTriP1())
    A = laplace.assemble(b)
    f = unit_load.assemble(b)
    A, f = femhip.enforce(A, f, D=m.boundary_nodes())
    u = spsolve(A.tocsc(), f)
    assert abs(u.max() - 0.073657185490792) < 1e-7


@pytest.mark.parametrize('case', [
    ('tri', 'tri1', 33, 0.2, None, 'laplace'),
    ('tri', 'tri2', 17, 0.2, None, 'mass'),
    ('tet', 'tet1', 9, 0.2, None, 'laplace'),
    ('tet', 'tet1', 9, 0.0, None, 'laplace'),
    ('tet', 'tet2', 6, 0.2, None, 'laplace'),
    ('hex', 'hex1', 6, 0.15, None, 'laplace'),
    ('hex', 'hex1', 5, 0.15, 5, 'mass'),
])
def test_local_entries_bitexact(case):
    """The exact-order kernel reproduces the oracle's COO data bit for bit."""
    import torch
    kind, el, n, a, io, form = case
    x = np.linspace(0, 1, n)
    m = {'tri': lambda: femhip.MeshTri.init_tensor(x, x),
         'tet': lambda: femhip.MeshTet.init_tensor(x, x, x),
         'hex': lambda: femhip.MeshHex.init_tensor(x, x, x)}[kind]()
    if a:
        m = m.jittered(a, 1.0 / (n - 1))
    b = femhip.Basis(m, ELEM[el](), intorder=io)
    c = b.device().local_bilinear(form)
    torch.cuda.synchronize()
    om = {'tri': O.MeshTri, 'tet': O.MeshTet, 'hex': O.MeshHex}[kind](m.p, m.t)
    oel = {'tri1': O.ElementTriP1, 'tri2': O.ElementTriP2, 'tet1': O.ElementTetP1,
           'tet2': O.ElementTetP2, 'hex1': O.ElementHex1}[el]()
    ob = O.CellBasis(om, oel, intorder=io)
    f = {'laplace': O.form_laplace, 'mass': O.form_mass}[form]
    _, data, _ = O.bilinear_coo(f, ob)
    got = c.cpu().numpy()
    assert got.shape == data.shape
    assert np.array_equal(got, data), np.abs(got - data).max()


def test_zero_jacobian_raises():
    p = np.array([[0., 0., 0.], [0., 0., 1.], [0., 1., 0.], [1., 0., 0.],
                  [0., 1., 1.], [1., 0., 1.], [1., 1., 0.], [1., 1., 1.]]).T
    p[:, 3] = p[:, 0]
    p[:, 5] = p[:, 1]
    p[:, 6] = p[:, 2]
    p[:, 7] = p[:, 4]   # flat hexahedron: x-extent zero
    m = femhip.MeshHex(p, np.arange(8, dtype=np.int32)[:, None])
    b = femhip.Basis(m, femhip.ElementHex1())
    with pytest.raises(Exception, match="Zero Jacobian"):
        laplace.assemble(b)


def test_enforce_on_device():
    import torch
    from femhip.utils import enforce_device
    x = np.linspace(0, 1, 33)
    m = femhip.MeshTri.init_tensor(x, x).jittered(0.2, 1 / 32)
    b = femhip.Basis(m, femhip.ElementTriP1())
    A = laplace.assemble_device(b)
    f = unit_load.assemble_device(b)
    D = m.boundary_nodes()
    xD = np.random.default_rng(0).standard_normal(b.N)
    Ah, fh = femhip.enforce(A.to_scipy(), f.cpu().numpy(), x=xD, D=D)
    enforce_device(A, f, x=xD, D=D)
    torch.cuda.synchronize()
    assert np.array_equal(A.data.cpu().numpy(), Ah.data)
    assert np.array_equal(f.cpu().numpy(), fh)


@pytest.mark.parametrize('kind,elem,n', [('tet', 'tet1', 13), ('tri', 'tri2', 17),
                                         ('hex', 'hex1', 6)])
def test_vector_laplace_vs_oracle(kind, elem, n):
    """vector_laplace on ElementVector(e), larger jittered meshes, first and
    second call (P1: the fused kernel for the scalar matrix), and a user form
    ddot(grad(u), grad(v)) recognised by the tracer."""
    from femhip.helpers import ddot, grad
    print('case', kind, elem, n, flush=True)
    x = np.linspace(0, 1, n)
    M = MESH[kind]
    m = (M.init_tensor(x, x) if kind == 'tri' else M.init_tensor(x, x, x)).jittered(
        0.15, 1.0 / (n - 1))
    b = femhip.Basis(m, femhip.ElementVector(ELEM[elem]()))
    om = {'tri': O.MeshTri, 'tet': O.MeshTet, 'hex': O.MeshHex}[kind](m.p, m.t)
    oe = {'tet1': O.ElementTetP1, 'tri2': O.ElementTriP2, 'hex1': O.ElementHex1}[elem]()
    R = O.assemble_named('vector_laplace', O.CellBasis(om, O.ElementVector(oe)))
    user = femhip.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v)))
    for A in (vector_laplace.assemble(b), vector_laplace.assemble(b), user.assemble(b)):
        assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
        assert np.abs(A.data - R.data).max() <= TOL * np.abs(R.data).max()


@pytest.mark.parametrize('kind,elem', [('tri', 'tri1'), ('tri', 'tri2'), ('tet', 'tet1'),
                                       ('tet', 'tet2')])
def test_single_element(kind, elem):
    """One element: laplace, mass and unit_load, twice (P1: the second call
    runs the fused kernel on a one-partition plan), against the oracle."""
    if kind == 'tri':
        p, t = np.array([[0.1, 1.3, 0.2], [0.0, 0.2, 0.9]]), np.array([[0], [1], [2]])
    else:
        p = np.array([[0.0, 1.1, 0.1, 0.2], [0.0, 0.1, 0.9, 0.1], [0.0, 0.2, 0.1, 1.2]])
        t = np.array([[0], [1], [2], [3]])
    m = MESH[kind](p, t)
    b = femhip.Basis(m, ELEM[elem]())
    om = {'tri': O.MeshTri, 'tet': O.MeshTet}[kind](m.p, m.t)
    ob = O.CellBasis(om, getattr(O, ELEM[elem].__name__)())
    for name, form in (('laplace', laplace), ('mass', mass)):
        R = O.assemble_named(name, ob)
        for _ in range(2):
            A = form.assemble(b)
            assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
            assert np.abs(A.data - R.data).max() <= TOL * np.abs(R.data).max()
    r = O.assemble_named('unit_load', ob)
    assert np.abs(unit_load.assemble(b) - r).max() <= TOL * np.abs(r).max()
