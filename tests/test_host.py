"""Host-side logic of the package against the oracle, and the C ABI
library's exports (CPU only; no compute call needs a GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import femhip
from femhip import quadrature as Q, trace
from femhip.models import laplace, mass, unit_load, kappa_laplace, linear_elasticity
from femhip.helpers import dot, grad
from oracle import skfem_np as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _meshes():
    x = np.linspace(0, 1, 5)
    y = np.linspace(0, 2, 4)
    return [
        (femhip.MeshTri.init_tensor(x, y), O.MeshTri.init_tensor(x, y)),
        (femhip.MeshTet.init_tensor(x, y, x), O.MeshTet.init_tensor(x, y, x)),
        (femhip.MeshHex.init_tensor(x, y, x), O.MeshHex.init_tensor(x, y, x)),
        (femhip.MeshTri(), O.MeshTri()),
        (femhip.MeshTet(), O.MeshTet()),
    ]


@pytest.mark.parametrize('i', range(5))
def test_mesh_matches_oracle(i):
    m, o = _meshes()[i]
    assert np.array_equal(m.p, o.p) and np.array_equal(m.t, o.t)
    assert np.array_equal(m.facets, o.facets)
    assert np.array_equal(m.t2f, o.t2f)
    if o.edge_idx is not None:
        assert np.array_equal(m.edges, o.edges)
        assert np.array_equal(m.t2e, o.t2e)
    assert np.array_equal(m.boundary_nodes(), o.boundary_nodes())
    assert np.array_equal(m.interior_nodes(), o.interior_nodes())


def test_jitter_matches_oracle():
    x = np.linspace(0, 1, 7)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1 / 6)
    o = O.jitter(O.MeshTet.init_tensor(x, x, x), 0.2, 1 / 6)
    assert np.array_equal(m.p, o.p)


ELS = [('tri', femhip.ElementTriP1, O.ElementTriP1),
       ('tri', femhip.ElementTriP2, O.ElementTriP2),
       ('tet', femhip.ElementTetP1, O.ElementTetP1),
       ('tet', femhip.ElementTetP2, O.ElementTetP2),
       ('hex', femhip.ElementHex1, O.ElementHex1),
       ('hex', femhip.ElementHex2, O.ElementHex2)]


@pytest.mark.parametrize('kind,E,OE', ELS)
@pytest.mark.parametrize('vector', [False, True])
def test_dofs_and_tables_match_oracle(kind, E, OE, vector):
    x = np.linspace(0, 1, 4)
    m = {'tri': lambda: femhip.MeshTri.init_tensor(x, x),
         'tet': lambda: femhip.MeshTet.init_tensor(x, x, x),
         'hex': lambda: femhip.MeshHex.init_tensor(x, x, x)}[kind]()
    om = {'tri': O.MeshTri, 'tet': O.MeshTet, 'hex': O.MeshHex}[kind](m.p, m.t)
    e, oe = E(), OE()
    if vector:
        e, oe = femhip.ElementVector(e), O.ElementVector(oe)
    b = femhip.Basis(m, e)
    ed, N = O.element_dofs(om, oe)
    assert np.array_equal(b.element_dofs, ed) and b.N == N
    oq = {'tri': O.quadrature_tri, 'tet': O.quadrature_tet, 'hex': O.quadrature_hex}[kind]
    X, W = oq(2 * oe.maxdeg)
    assert np.array_equal(b.X, X) and np.array_equal(b.W, W)
    phi, dphi, _ = b.tables()
    base = oe.elem if vector else oe
    for i in range(base.nbfun):
        v, g = base.lbasis(X, i)
        assert np.array_equal(phi[i], np.broadcast_to(v, phi[i].shape))
        assert np.array_equal(dphi[i], np.broadcast_to(g, dphi[i].shape))


@pytest.mark.parametrize('kind,order', [('tri', o) for o in (1, 2, 3, 4)]
                         + [('tet', o) for o in (1, 2, 3, 4)]
                         + [('hex', o) for o in (1, 3, 5, 7, 9, 11, 13)])
def test_quadrature_matches_oracle(kind, order):
    X, W = Q.get_quadrature(kind, order)
    oX, oW = {'tri': O.quadrature_tri, 'tet': O.quadrature_tet,
              'hex': O.quadrature_hex}[kind](order)
    assert np.array_equal(X, oX) and np.array_equal(W, oW)


def test_enforce_matches_oracle():
    x = np.linspace(0, 1, 9)
    o = O.MeshTri.init_tensor(x, x)
    ob = O.CellBasis(o, O.ElementTriP1())
    A = O.bilinear_assemble(O.form_laplace, ob)
    f = O.linear_assemble(O.form_unit_load, ob)
    D = o.boundary_nodes()
    A1, f1 = femhip.enforce(A, f, D=D)
    A2, f2 = O.enforce(A, f, D=D)
    assert (A1 != A2).nnz == 0 and np.array_equal(f1, f2)
    A3 = femhip.enforce(A, I=o.interior_nodes())
    assert (A3 != A2).nnz == 0


def test_form_recognition():
    @femhip.BilinearForm
    def lap(u, v, _):
        return dot(grad(u), grad(v))

    @femhip.BilinearForm
    def m2(u, v, w):
        return 2.0 * u * v

    @femhip.BilinearForm
    def kl(u, v, w):
        return dot(w.sigma * grad(u), grad(v))

    @femhip.LinearForm
    def rhs(v, _):
        return 1.0 * v

    assert lap.spec() == ('laplace', 1.0, {})
    assert m2.spec() == ('mass', 2.0, {})
    assert kl.spec() == ('kappa_laplace', 1.0, {'coef': 'sigma'})
    assert rhs.spec() == ('unit_load', 1.0, {})
    assert laplace.spec()[0] == 'laplace' and mass.spec()[0] == 'mass'
    assert unit_load.spec()[0] == 'unit_load'
    assert kappa_laplace.spec()[0] == 'kappa_laplace'
    assert linear_elasticity(1.0, 2.0).spec() == (
        'linear_elasticity', 1.0, {'Lambda': 1.0, 'Mu': 2.0})

    @femhip.BilinearForm
    def unknown(u, v, w):
        return u * grad(v)[0]

    with pytest.raises(NotImplementedError):
        unknown.spec()


def test_library_exports_every_declared_symbol():
    """The C ABI library loads without a GPU and exports every function
    include/femhip.h declares."""
    from femhip import _lib
    header = open(os.path.join(ROOT, 'include', 'femhip.h')).read()
    names = sorted(set(re.findall(r'\b(femhip_[a-z0-9_]+)\s*\(', header)))
    assert names, "no declarations found"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    for n in names:
        assert n in _lib.SIGNATURES, "no ctypes signature for %s" % n
    lib2 = _lib.load()
    assert lib2.femhip_version() >= 1
    # argument validation needs no device: a NULL geometry is rejected
    assert lib2.femhip_geometry_size(None) == -1


def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    b = femhip.Basis(femhip.MeshTri(), femhip.ElementTriP1())
    with pytest.raises(RuntimeError):
        laplace.assemble(b)


def test_laplace_rows_masks_the_slots_the_laplace_pattern_lacks():
    """Per plan row: the Laplace row start and a bit per mass slot absent
    from the Laplace row (the fused Laplace + mass pass compacts with it)."""
    from femhip.fused import laplace_rows
    mi = np.array([0, 3, 6, 8])
    mx = np.array([0, 1, 2, 0, 1, 2, 1, 2])
    li = np.array([0, 2, 5, 7])
    lx = np.array([0, 2, 0, 1, 2, 1, 2])          # row 0 lacks column 1
    plan_rows = np.array([3, 0, 0, 0, 6, 0], dtype=np.int32)   # rows 1, 0, 2
    out = laplace_rows(plan_rows, mi, mx, li, lx).reshape(-1, 2)
    assert out[:, 0].tolist() == [2, 0, 5]
    assert out[:, 1].tolist() == [0, 0b010, 0]
