"""Coefficient forms and the default form parameters on the device, against
the reference's own outputs (tests/golden/facet_*.npz, coef_*.npz):

* Robin ``w['alpha'] * u * v`` and Neumann ``w['g'] * v`` written as user
  forms (recognised by the tracer), on a FacetBasis and on a CellBasis:
  pattern bit-exact, values within 1e-12 of max|A| (max|b|);
* a DOF vector passed as the coefficient (interpolated as the reference
  does), and the facet interpolation itself (bit-exact);
* the quadrature points (bit-exact) and the mesh parameter (bit-exact
  where NumPy's power takes sqrt or a copy, within 2 ulp for the exponent
  1/3, which both sides evaluate with their own ``pow``).
"""
import glob
import json
import os

import numpy as np
import pytest

import femhip
from femhip import BilinearForm, LinearForm
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
FACET = sorted(glob.glob(os.path.join(GOLD, 'facet_*.npz')))
CELL = sorted(glob.glob(os.path.join(GOLD, 'coef_*.npz')))
FELEM = {'tri1': femhip.ElementTriP1, 'tri2': femhip.ElementTriP2,
         'tet1': femhip.ElementTetP1, 'tet2': femhip.ElementTetP2,
         'hex1': femhip.ElementHex1}


@BilinearForm
def robin(u, v, w):
    return w['alpha'] * u * v


@LinearForm
def neumann(v, w):
    return w['g'] * v


def _close(a, ref, tol=1e-12):
    return a.shape == ref.shape and (a.size == 0 or
                                     np.abs(a - ref).max() <= tol * np.abs(ref).max())


def _check_csr(A, z, key):
    assert np.array_equal(A.indptr, z[key + '_indptr'])
    assert np.array_equal(A.indices, z[key + '_indices'])
    assert _close(A.data, z[key + '_data'])


@pytest.mark.parametrize('fn', FACET, ids=lambda f: os.path.basename(f)[6:-4])
def test_facet_coefficient_forms(fn):
    z = np.load(fn)
    meta = json.loads(str(z['meta']))
    print('case', os.path.basename(fn), flush=True)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet}[meta['kind']]
    m = M(z['p'], z['t'])
    facets = None if meta['facets'] == 'boundary' else z['find']
    fb = femhip.FacetBasis(m, FELEM[meta['element']](), facets=facets, side=meta['side'])
    x = fb.global_coordinates()
    assert np.array_equal(np.asarray(x), z['x'])
    assert np.array_equal(np.asarray(fb.mesh_parameters()), z['h'])
    _check_csr(robin.assemble(fb, alpha=z['alpha']), z, 'robin')
    assert _close(neumann.assemble(fb, g=z['g']), z['neumann_b'])
    ip = fb.interpolate(z['wdof'])
    assert np.array_equal(ip.value, z['interp_value'])
    assert np.array_equal(ip.grad, z['interp_grad'])
    assert _close(neumann.assemble(fb, g=z['wdof']), z['neumann_dof_b'])
    # the coefficient computed from w.x of the device basis gives the same
    alpha = 1.0 + x[0] ** 2 + 0.5 * x[1]
    _check_csr(robin.assemble(fb, alpha=alpha), z, 'robin')
    assert set(fb.default_parameters()) == {'x', 'h', 'n'}


@pytest.mark.parametrize('fn', CELL, ids=lambda f: os.path.basename(f)[5:-4])
def test_cell_coefficient_forms(fn):
    z = np.load(fn)
    meta = json.loads(str(z['meta']))
    print('case', os.path.basename(fn), flush=True)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet, 'hex': femhip.MeshHex}[meta['kind']]
    b = femhip.Basis(M(z['p'], z['t']), FELEM[meta['element']]())
    assert np.array_equal(np.asarray(b.global_coordinates()), z['x'])
    h = np.asarray(b.mesh_parameters())
    if meta['kind'] == 'tri':
        assert np.array_equal(h, z['h'])
    else:
        assert np.allclose(h, z['h'], rtol=4.5e-16, atol=0)
    _check_csr(robin.assemble(b, alpha=z['alpha']), z, 'robin')
    assert _close(neumann.assemble(b, g=z['g']), z['neumann_b'])
    assert _close(neumann.assemble(b, g=z['wdof']), z['neumann_dof_b'])
    assert set(b.default_parameters()) == {'x', 'h'}


@pytest.mark.parametrize('kind,elem,n', [('tri', 'tri2', 33), ('tet', 'tet2', 9)])
def test_facet_robin_vs_oracle_larger(kind, elem, n):
    """Larger jittered meshes against the oracle, per-facet coefficient
    (nf, 1) and per-point coefficient; local matrices bit-identical."""
    print('case', kind, elem, n, flush=True)
    x = np.linspace(0, 1, n)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet}[kind]
    m = (M.init_tensor(x, x) if kind == 'tri' else M.init_tensor(x, x, x)).jittered(
        0.2, 1.0 / (n - 1))
    fb = femhip.FacetBasis(m, FELEM[elem]())
    ob = O.FacetBasis({'tri': O.MeshTri, 'tet': O.MeshTet}[kind](m.p, m.t),
                      getattr(O, FELEM[elem].__name__)())
    rng = np.random.default_rng(5)
    a_pt = rng.lognormal(0, 1, ob.dx.shape)
    a_f = rng.lognormal(0, 1, (ob.nelems, 1))
    for a in (a_pt, a_f):
        A = robin.assemble(fb, alpha=a)
        R = O.bilinear_assemble(O.form_coef_mass, ob, c=a)
        assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
        assert _close(A.data, R.data)
        assert _close(neumann.assemble(fb, g=a), O.linear_assemble(O.form_coef_load, ob, c=a))
    loc = robin.elemental(fb, alpha=a_pt)
    nb = ob.Nbfun
    ref = np.empty((ob.nelems, nb, nb))
    for j in range(nb):
        for i in range(nb):
            ref[:, i, j] = np.sum(O.form_coef_mass(ob.basis[j], ob.basis[i], {'c': a_pt})
                                  * ob.dx, axis=1)
    assert np.array_equal(loc, ref)
