"""FacetBasis on the CPU: the oracle's restatement against the reference's
own outputs (tests/golden/facet_*.npz, written by make_fixtures.py) and
against closed forms, and the package's host-side facet data (facets,
owning elements, DOFs, quadrature) against the same fixtures."""
import glob
import json
import os

import numpy as np
import pytest

import femhip
from oracle import skfem_np as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
FIXTURES = sorted(glob.glob(os.path.join(GOLD, 'facet_*.npz')))
OELEM = {'tri1': O.ElementTriP1, 'tri2': O.ElementTriP2,
         'tet1': O.ElementTetP1, 'tet2': O.ElementTetP2}
FELEM = {'tri1': femhip.ElementTriP1, 'tri2': femhip.ElementTriP2,
         'tet1': femhip.ElementTetP1, 'tet2': femhip.ElementTetP2}


def _load(fn):
    z = np.load(fn)
    return z, json.loads(str(z['meta']))


def _ofacet(z, meta, **kw):
    m = {'tri': O.MeshTri, 'tet': O.MeshTet}[meta['kind']](z['p'], z['t'])
    facets = None if meta['facets'] == 'boundary' else z['find']
    return O.FacetBasis(m, OELEM[meta['element']](), facets=facets, side=meta['side'], **kw)


def test_fixtures_present():
    assert len(FIXTURES) == 7


@pytest.mark.parametrize('fn', FIXTURES, ids=lambda f: os.path.basename(f)[6:-4])
def test_oracle_matches_reference(fn):
    """Bit-identical facet data, CSR pattern, mass values and load vector."""
    z, meta = _load(fn)
    fb = _ofacet(z, meta)
    for k in ('find', 'tind', 'element_dofs', 'X', 'W', 'dx', 'normals'):
        assert np.array_equal(getattr(fb, k), z[k]), k
    A = O.bilinear_assemble(O.form_mass, fb)
    assert np.array_equal(A.indptr, z['mass_indptr'])
    assert np.array_equal(A.indices, z['mass_indices'])
    assert np.array_equal(A.data, z['mass_data'])
    b = O.linear_assemble(O.form_unit_load, fb)
    assert np.array_equal(b, z['unit_load_b'])


def _nodal(m, elem, f):
    """Nodal interpolant of f(x) for P1/P2 (vertices, then edge midpoints)."""
    b = O.CellBasis(m, elem)
    ed = b.element_dofs
    locs = np.zeros((m.p.shape[0], b.N))
    nv = m.t.shape[0]
    locs[:, ed[:nv].ravel()] = m.p[:, m.t].reshape(m.p.shape[0], -1)
    if ed.shape[0] > nv:
        pairs = {2: ((0, 1), (1, 2), (0, 2)),
                 3: ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))}[m.p.shape[0]]
        for k, (a, c) in enumerate(pairs):
            mid = 0.5 * (m.p[:, m.t[a]] + m.p[:, m.t[c]])
            locs[:, ed[nv + k]] = mid
    return f(locs)


@pytest.mark.parametrize('kind,elem,n', [('tri', 'tri1', 9), ('tri', 'tri2', 5),
                                         ('tet', 'tet1', 4), ('tet', 'tet2', 3)])
def test_oracle_boundary_integrals(kind, elem, n):
    """On the unit square / cube boundary: sum(load) = |dOmega|,
    1^T M 1 = |dOmega|, x^T M 1 = int x ds (P1 and P2 reproduce x), and
    for P2 (x^2)^T M 1 = int x^2 ds -- on a jittered mesh (interior nodes
    only move, so the boundary stays exact)."""
    x = np.linspace(0, 1, n)
    m = (O.MeshTri.init_tensor(x, x) if kind == 'tri' else O.MeshTet.init_tensor(x, x, x))
    m = O.jitter(m, 0.2, 1.0 / (n - 1))
    fb = O.FacetBasis(m, OELEM[elem]())
    A = O.bilinear_assemble(O.form_mass, fb)
    b = O.linear_assemble(O.form_unit_load, fb)
    area = 4.0 if kind == 'tri' else 6.0
    one = np.ones(fb.N)
    assert abs(b.sum() - area) < 1e-13
    assert abs(one @ A @ one - area) < 1e-13
    u = _nodal(m, OELEM[elem](), lambda X: X[0])
    int_x = 2.0 if kind == 'tri' else 3.0          # x=1 side + sides where x varies
    assert abs(u @ A @ one - int_x) < 1e-13
    assert abs(b @ u - int_x) < 1e-13
    if elem.endswith('2'):
        u2 = _nodal(m, OELEM[elem](), lambda X: X[0] ** 2)
        int_x2 = 1.0 + (2.0 if kind == 'tri' else 4.0) / 3.0
        assert abs(u2 @ A @ one - int_x2) < 1e-13
        # int x^2 * x ds = int x^3 ds (degree 4 product, exact at order 4)
        int_x3 = 1.0 + (2.0 if kind == 'tri' else 4.0) / 4.0
        assert abs(u2 @ A @ u - int_x3) < 1e-13


def test_oracle_p1_segment_mass():
    """One boundary segment of length L: the P1 facet mass is L/6 [[2,1],[1,2]]."""
    p = np.array([[0., 2., 0.3], [0., 0., 1.1]])
    t = np.array([[0], [1], [2]])
    m = O.MeshTri(p, t)
    fb = O.FacetBasis(m, O.ElementTriP1(), facets=np.array([0]))
    assert np.array_equal(m.facets[:, 0], [0, 1])
    A = O.bilinear_assemble(O.form_mass, fb).toarray()
    ref = np.zeros((3, 3))
    ref[:2, :2] = 2.0 / 6.0 * np.array([[2., 1.], [1., 2.]])
    assert np.allclose(A, ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize('kind,n', [('tri', 5), ('tet', 3)])
def test_oracle_normals_outward(kind, n):
    """Boundary normals of the unit square / cube are the outward axis
    directions."""
    x = np.linspace(0, 1, n)
    m = (O.MeshTri.init_tensor(x, x) if kind == 'tri' else O.MeshTet.init_tensor(x, x, x))
    m = O.jitter(m, 0.2, 1.0 / (n - 1))
    fb = O.FacetBasis(m, (O.ElementTriP1 if kind == 'tri' else O.ElementTetP1)())
    mid = m.p[:, m.facets[:, fb.find]].mean(axis=1)
    nrm = fb.normals[:, :, 0]
    ax = np.argmax(np.abs(nrm), axis=0)
    sgn = np.sign(nrm[ax, np.arange(len(ax))])
    assert np.allclose(np.abs(nrm).max(axis=0), 1.0, atol=1e-15)
    assert np.all(sgn == np.where(mid[ax, np.arange(len(ax))] > 0.5, 1, -1))
    assert np.allclose(np.sum(nrm ** 2, axis=0), 1.0, atol=1e-15)


@pytest.mark.parametrize('fn', FIXTURES, ids=lambda f: os.path.basename(f)[6:-4])
def test_host_facet_data(fn):
    """The package's FacetBasis selects the same facets, owning elements,
    DOFs and quadrature as the reference."""
    z, meta = _load(fn)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet}[meta['kind']]
    m = M(z['p'], z['t'])
    facets = None if meta['facets'] == 'boundary' else z['find']
    fb = femhip.FacetBasis(m, FELEM[meta['element']](), facets=facets, side=meta['side'])
    assert np.array_equal(fb.find, z['find'])
    assert np.array_equal(fb.tind, z['tind'])
    assert np.array_equal(fb.element_dofs, z['element_dofs'])
    assert np.array_equal(fb.X, z['X']) and np.array_equal(fb.W, z['W'])
    assert int(fb.N) == int(z['N'])


def test_host_facet_selection():
    x = np.linspace(0, 1, 5)
    m = femhip.MeshTri.init_tensor(x, x)
    om = O.MeshTri(m.p, m.t)
    assert np.array_equal(m.f2t, om.f2t)
    left = m.normalize_facets(lambda X: X[0] < 1e-12)
    assert len(left) == 4
    both = m.normalize_facets([left, 0])
    assert np.array_equal(both, np.unique(np.concatenate([left, [0]])))
    assert np.array_equal(m.normalize_facets(None), m.boundary_facets())
    with pytest.raises(ValueError):
        m.normalize_facets('left')
    with pytest.raises(ValueError):
        femhip.FacetBasis(m, femhip.ElementTriP1(), side=1)   # boundary: no side 1
    with pytest.raises(NotImplementedError):
        femhip.FacetBasis(femhip.MeshHex.init_tensor(x, x, x), femhip.ElementHex1())


@pytest.mark.parametrize('fn', FIXTURES, ids=lambda f: os.path.basename(f)[6:-4])
def test_oracle_coefficient_forms_match_reference(fn):
    """Facet points, mesh parameter, Robin matrix alpha*u*v, Neumann vector
    g*v, a DOF-vector coefficient (interpolated on the facets) and the
    interpolation itself: bit-identical to the reference's."""
    z, meta = _load(fn)
    fb = _ofacet(z, meta)
    assert np.array_equal(fb.x, z['x']) and np.array_equal(fb.h, z['h'])
    R = O.bilinear_assemble(O.form_coef_mass, fb, c=z['alpha'])
    for k in ('indptr', 'indices', 'data'):
        assert np.array_equal(getattr(R, k), z['robin_' + k]), k
    assert np.array_equal(O.linear_assemble(O.form_coef_load, fb, c=z['g']), z['neumann_b'])
    v, g = O.interpolate(fb, z['wdof'])
    assert np.array_equal(v, z['interp_value']) and np.array_equal(g, z['interp_grad'])
    assert np.array_equal(O.linear_assemble(O.form_coef_load, fb, c=v), z['neumann_dof_b'])


@pytest.mark.parametrize('kind,elem,n', [('tri', 'tri1', 9), ('tri', 'tri2', 5),
                                         ('tet', 'tet1', 4), ('tet', 'tet2', 3)])
def test_oracle_robin_closed_form(kind, elem, n):
    """1^T R 1 = int alpha ds for alpha = 1 + x^2 + y/2 on the unit square /
    cube boundary (exact: the facet rules integrate quadratics), and the
    facet points lie on their facets' planes."""
    x = np.linspace(0, 1, n)
    m = (O.MeshTri.init_tensor(x, x) if kind == 'tri' else O.MeshTet.init_tensor(x, x, x))
    m = O.jitter(m, 0.2, 1.0 / (n - 1))
    fb = O.FacetBasis(m, OELEM[elem]())
    alpha = 1.0 + fb.x[0] ** 2 + 0.5 * fb.x[1]
    R = O.bilinear_assemble(O.form_coef_mass, fb, c=alpha)
    one = np.ones(fb.N)
    if kind == 'tri':
        exact = 1.25 + 2.25 + 4.0 / 3.0 + 11.0 / 6.0
    else:
        # faces x=0, x=1 (alpha in y), y=0, y=1 (in x), z=0, z=1 (in x, y)
        exact = 1.25 + 2.25 + 4.0 / 3.0 + 11.0 / 6.0 + 2 * (1.0 + 1.0 / 3.0 + 0.25)
    assert abs(one @ R @ one - exact) < 1e-13
    on_plane = np.min(np.abs(np.stack([fb.x, 1.0 - fb.x])), axis=(0, 1))
    assert on_plane.max() < 1e-15
