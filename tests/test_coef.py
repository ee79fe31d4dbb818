"""Coefficient forms on a CellBasis on the CPU: the oracle against the
reference's own outputs (tests/golden/coef_*.npz) and closed forms, and the
tracer's recognition of ``w.c * u * v`` and ``w.c * v``."""
import glob
import json
import os

import numpy as np
import pytest

from femhip.trace import identify
from oracle import skfem_np as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
FIXTURES = sorted(glob.glob(os.path.join(GOLD, 'coef_*.npz')))
OELEM = {'tri2': O.ElementTriP2, 'tet1': O.ElementTetP1, 'tet2': O.ElementTetP2,
         'hex1': O.ElementHex1}


def _obasis(z, meta):
    m = {'tri': O.MeshTri, 'tet': O.MeshTet, 'hex': O.MeshHex}[meta['kind']](z['p'], z['t'])
    return O.CellBasis(m, OELEM[meta['element']]())


def test_fixtures_present():
    assert len(FIXTURES) == 4


@pytest.mark.parametrize('fn', FIXTURES, ids=lambda f: os.path.basename(f)[5:-4])
def test_oracle_matches_reference(fn):
    z = np.load(fn)
    meta = json.loads(str(z['meta']))
    b = _obasis(z, meta)
    assert np.array_equal(b.global_coordinates(), z['x'])
    assert np.array_equal(b.mesh_parameters(), z['h'])
    R = O.bilinear_assemble(O.form_coef_mass, b, c=z['alpha'])
    for k in ('indptr', 'indices', 'data'):
        assert np.array_equal(getattr(R, k), z['robin_' + k]), k
    assert np.array_equal(O.linear_assemble(O.form_coef_load, b, c=z['g']), z['neumann_b'])
    v, _ = O.interpolate(b, z['wdof'])
    ref = z['neumann_dof_b']
    assert np.abs(O.linear_assemble(O.form_coef_load, b, c=v) - ref).max() <= 1e-15 * np.abs(ref).max()


@pytest.mark.parametrize('kind,elem,n', [('tri', 'tri2', 5), ('tet', 'tet2', 3),
                                         ('hex', 'hex1', 3)])
def test_oracle_closed_forms(kind, elem, n):
    """int alpha dx = 1^T R 1 for alpha = 1 + x^2 + y/2 over the unit square /
    cube (1 + 1/3 + 1/4), int g v for g = 1 sums to the volume, and the
    quadrature points of an undistorted mesh reproduce sum(W |det| x) = 1/2."""
    x = np.linspace(0, 1, n)
    M = {'tri': O.MeshTri, 'tet': O.MeshTet, 'hex': O.MeshHex}[kind]
    m = M.init_tensor(x, x) if kind == 'tri' else M.init_tensor(x, x, x)
    b = O.CellBasis(m, OELEM[elem]())
    X = b.global_coordinates()
    alpha = 1.0 + X[0] ** 2 + 0.5 * X[1]
    R = O.bilinear_assemble(O.form_coef_mass, b, c=alpha)
    one = np.ones(b.N)
    assert abs(one @ R @ one - (1.0 + 1.0 / 3.0 + 0.25)) < 1e-13
    assert abs(O.linear_assemble(O.form_coef_load, b, c=np.ones_like(alpha)).sum() - 1.0) < 1e-13
    assert abs(np.sum(b.dx * X[0]) - 0.5) < 1e-14
    h = b.mesh_parameters()
    vol = np.abs(b.mapping.detDF(b.X))
    assert np.allclose(h ** m.p.shape[0], vol, rtol=1e-14, atol=0)


def test_tracer_coefficient_forms():
    assert identify(lambda u, v, w: w['alpha'] * u * v, 3) == ('coef_mass', 1.0, {'coef': 'alpha'})
    assert identify(lambda u, v, w: 2.0 * u * v * w.a, 3) == ('coef_mass', 2.0, {'coef': 'a'})
    assert identify(lambda v, w: w.g * v, 2) == ('coef_load', 1.0, {'coef': 'g'})
    assert identify(lambda v, w: v * w['g'] / 4, 2) == ('coef_load', 0.25, {'coef': 'g'})
    with pytest.raises(NotImplementedError):
        identify(lambda u, v, w: w.a * w.b * u * v, 3)
    with pytest.raises(NotImplementedError):
        identify(lambda v, w: w.a * w.b * v, 2)
