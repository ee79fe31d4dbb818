"""Boundary forms on a FacetBasis on the device: facet mass (Robin) and
facet load (Neumann) against the reference's own outputs
(tests/golden/facet_*.npz) and against the oracle on larger jittered meshes;
pattern bit-exact, values within 1e-12 of max|A| (the load within 1e-12 of
max|b|), normals within 1e-15."""
import glob
import json
import os

import numpy as np
import pytest

import femhip
from femhip.models import mass, unit_load
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
FIXTURES = sorted(glob.glob(os.path.join(GOLD, 'facet_*.npz')))
FELEM = {'tri1': femhip.ElementTriP1, 'tri2': femhip.ElementTriP2,
         'tet1': femhip.ElementTetP1, 'tet2': femhip.ElementTetP2}
OELEM = {'tri1': O.ElementTriP1, 'tri2': O.ElementTriP2,
         'tet1': O.ElementTetP1, 'tet2': O.ElementTetP2}


def _close(a, ref, tol=1e-12):
    scale = max(np.abs(ref).max(), 1e-300) if ref.size else 1.0
    return a.shape == ref.shape and (a.size == 0 or np.abs(a - ref).max() <= tol * scale)


@pytest.mark.parametrize('fn', FIXTURES, ids=lambda f: os.path.basename(f)[6:-4])
def test_facet_forms_vs_reference(fn):
    z = np.load(fn)
    meta = json.loads(str(z['meta']))
    print('case', os.path.basename(fn), flush=True)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet}[meta['kind']]
    m = M(z['p'], z['t'])
    facets = None if meta['facets'] == 'boundary' else z['find']
    fb = femhip.FacetBasis(m, FELEM[meta['element']](), facets=facets, side=meta['side'])
    A = mass.assemble(fb)
    assert np.array_equal(A.indptr, z['mass_indptr'])
    assert np.array_equal(A.indices, z['mass_indices'])
    assert _close(A.data, z['mass_data'])
    b = unit_load.assemble(fb)
    assert _close(b, z['unit_load_b'])
    assert np.allclose(fb.normals.value, z['normals'], rtol=0, atol=1e-15)


@pytest.mark.parametrize('kind,elem,n,side', [
    ('tri', 'tri1', 65, 0), ('tri', 'tri2', 33, 0), ('tet', 'tet1', 13, 0),
    ('tet', 'tet2', 9, 0), ('tet', 'tet2', 9, 1), ('tri', 'tri2', 33, 1),
])
def test_facet_forms_vs_oracle(kind, elem, n, side):
    """Larger jittered meshes against the oracle; side 1 on the interior
    facets of the grid plane x = 1/2 (selected on the undistorted mesh)."""
    print('case', kind, elem, n, side, flush=True)
    x = np.linspace(0, 1, n)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet}[kind]
    m0 = M.init_tensor(x, x) if kind == 'tri' else M.init_tensor(x, x, x)
    m = m0.jittered(0.2, 1.0 / (n - 1))
    facets = None
    if side == 1:
        facets = m0.facets_satisfying(lambda X: np.abs(X[0] - 0.5) < 1e-9)
        assert len(facets) > 0
    fb = femhip.FacetBasis(m, FELEM[elem](), facets=facets, side=side)
    om = {'tri': O.MeshTri, 'tet': O.MeshTet}[kind](m.p, m.t)
    ob = O.FacetBasis(om, OELEM[elem](), facets=facets, side=side)
    A = mass.assemble(fb)
    R = O.bilinear_assemble(O.form_mass, ob)
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert _close(A.data, R.data)
    assert _close(unit_load.assemble(fb), O.linear_assemble(O.form_unit_load, ob))
    assert np.allclose(fb.normals.value, ob.normals, rtol=0, atol=1e-15)


def test_facet_elemental_and_repeat():
    """Repeated assembly reuses the cached pattern and gives the same bits;
    the local matrices equal the oracle's per-facet sums."""
    x = np.linspace(0, 1, 17)
    m = femhip.MeshTri.init_tensor(x, x).jittered(0.2, 1 / 16)
    fb = femhip.FacetBasis(m, femhip.ElementTriP2())
    A1, A2 = mass.assemble(fb), mass.assemble(fb)
    assert np.array_equal(A1.data, A2.data) and np.array_equal(A1.indices, A2.indices)
    loc = mass.elemental(fb)
    ob = O.FacetBasis(O.MeshTri(m.p, m.t), O.ElementTriP2())
    nb = ob.Nbfun
    ref = np.empty((ob.nelems, nb, nb))
    for j in range(nb):
        for i in range(nb):
            ref[:, i, j] = np.sum(O.form_mass(ob.basis[j], ob.basis[i], {}) * ob.dx, axis=1)
    assert np.array_equal(loc, ref)


def test_facet_empty_set():
    x = np.linspace(0, 1, 5)
    m = femhip.MeshTri.init_tensor(x, x)
    fb = femhip.FacetBasis(m, femhip.ElementTriP1(), facets=np.zeros(0, dtype=np.int32))
    A = mass.assemble(fb)
    assert A.shape == (fb.N, fb.N) and A.nnz == 0
    assert np.array_equal(unit_load.assemble(fb), np.zeros(fb.N))
    assert fb.normals.value.shape == (2, 0, fb.W.shape[0])
    with pytest.raises(NotImplementedError):
        femhip.models.laplace.assemble(fb)


def test_robin_poisson_solve():
    """-Laplace u = 1 with u + du/dn = 0 on the boundary (Robin): the system
    assembled on the device (cell Laplace + facet mass) and the oracle's
    give the same matrix and, solved by SciPy, the same solution."""
    from femhip.models import laplace
    from scipy.sparse.linalg import spsolve
    x = np.linspace(0, 1, 33)
    m = femhip.MeshTri.init_tensor(x, x).jittered(0.2, 1 / 32)
    cb = femhip.Basis(m, femhip.ElementTriP2())
    fb = femhip.FacetBasis(m, femhip.ElementTriP2())
    K = laplace.assemble(cb) + mass.assemble(fb)
    f = unit_load.assemble(cb)
    om = O.MeshTri(m.p, m.t)
    oc, of = O.CellBasis(om, O.ElementTriP2()), O.FacetBasis(om, O.ElementTriP2())
    Kr = O.assemble_named('laplace', oc) + O.bilinear_assemble(O.form_mass, of)
    fr = O.assemble_named('unit_load', oc)
    assert np.abs((K - Kr).data).max() <= 1e-12 * np.abs(Kr.data).max()
    u = spsolve(K.tocsc(), f)
    ur = spsolve(Kr.tocsc(), fr)
    assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max()
