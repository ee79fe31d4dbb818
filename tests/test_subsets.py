"""Element subsets (``CellBasis(elements=...)``) and sums over several bases
on the CPU: the element criteria of the host, and the oracle's subset
assembly pinned by additivity -- the assembly over the two halves of an
element partition, concatenated as COO entries, is the assembly over all
elements (same pattern, values to rounding) -- and by the full index set
giving the full assembly bit for bit."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix

import femhip
from femhip.basis import normalize_elements, ElementSubset
from oracle import skfem_np as O


def _tri(n=9):
    x = np.linspace(0, 1, n)
    return femhip.MeshTri.init_tensor(x, x)


def test_normalize_elements_criteria():
    m = _tri()
    E = m.nelements
    assert np.array_equal(normalize_elements(m, True), np.arange(E))
    assert np.array_equal(normalize_elements(m, 3), [3])
    idx = np.array([5, 1, 9])
    assert np.array_equal(normalize_elements(m, idx), idx)      # kept as given
    mask = np.zeros(E, bool)
    mask[[2, 7]] = True
    assert np.array_equal(normalize_elements(m, mask), [2, 7])
    c = m.p[:, m.t].mean(axis=1)
    left = normalize_elements(m, lambda x: x[0] < 0.5)
    assert np.array_equal(left, np.flatnonzero(c[0] < 0.5))
    u = normalize_elements(m, [np.array([4, 2]), 2, 11])
    assert np.array_equal(u, [2, 4, 11])
    with pytest.raises(ValueError):
        normalize_elements(m, np.zeros(E + 1, bool))
    with pytest.raises(NotImplementedError):
        normalize_elements(m, 'left')


def test_subset_basis_host_side():
    m = _tri()
    tind = np.flatnonzero(m.p[0, m.t].mean(axis=0) < 0.5)
    b = femhip.CellBasis(m, femhip.ElementTriP2(), elements=tind)
    full = femhip.CellBasis(m, femhip.ElementTriP2())
    assert b.N == full.N and b.nelems == len(tind)
    assert np.array_equal(b.element_dofs, full.element_dofs[:, tind])
    assert isinstance(b.cell_mesh, ElementSubset)
    assert np.array_equal(b.cell_mesh.t, m.t[:, tind]) and b.cell_mesh.p is m.p
    assert b.cell_mesh.nelements == len(tind) and b.cell_mesh.kind == 'tri'
    assert np.array_equal(b.with_element(femhip.ElementTriP1()).tind, tind)
    assert full.tind is None and full.cell_mesh is m


def _omesh(kind, n, a):
    x = np.linspace(0, 1, n)
    m = O.MeshTri.init_tensor(x, x) if kind == 'tri' else O.MeshTet.init_tensor(x, x, x)
    return O.jitter(m, a, 1.0 / (n - 1)) if a else m


def _coo_csr(parts, shape):
    r = np.concatenate([p[0][0] for p in parts])
    c = np.concatenate([p[0][1] for p in parts])
    d = np.concatenate([p[1] for p in parts])
    K = coo_matrix((d, (r, c)), shape=shape)
    K.eliminate_zeros()
    return K.tocsr()


@pytest.mark.parametrize('kind,n,el,form', [
    ('tri', 9, 'ElementTriP1', 'laplace'), ('tri', 7, 'ElementTriP2', 'mass'),
    ('tet', 4, 'ElementTetP2', 'laplace'), ('tet', 5, 'ElementTetP1', 'mass'),
])
def test_oracle_subsets_are_additive(kind, n, el, form):
    m = _omesh(kind, n, 0.2)
    f = {'laplace': O.form_laplace, 'mass': O.form_mass}[form]
    full = O.CellBasis(m, getattr(O, el)())
    A = O.bilinear_assemble(f, full)
    E = m.nelements
    rng = np.random.default_rng(1)
    sel = rng.random(E) < 0.4
    t0, t1 = np.flatnonzero(sel), np.flatnonzero(~sel)
    b0 = O.CellBasis(m, getattr(O, el)(), elements=t0)
    b1 = O.CellBasis(m, getattr(O, el)(), elements=t1)
    assert b0.N == full.N and b0.nelems == len(t0)
    c0 = O.bilinear_coo(f, b0)
    c1 = O.bilinear_coo(f, b1)
    S = _coo_csr([c0[:2], c1[:2]], A.shape)
    assert np.array_equal(S.indptr, A.indptr) and np.array_equal(S.indices, A.indices)
    assert np.abs(S.data - A.data).max() <= 1e-13 * np.abs(A.data).max()
    # the full index set in element order is the unrestricted basis
    ball = O.CellBasis(m, getattr(O, el)(), elements=np.arange(E))
    B = O.bilinear_assemble(f, ball)
    assert np.array_equal(B.indices, A.indices) and np.array_equal(B.data, A.data)
    # load vectors add up the same way
    l0 = O.linear_assemble(O.form_unit_load, b0)
    l1 = O.linear_assemble(O.form_unit_load, b1)
    lf = O.linear_assemble(O.form_unit_load, full)
    assert np.abs(l0 + l1 - lf).max() <= 1e-14 * np.abs(lf).max()


def test_oracle_vector_interpolation_is_componentwise():
    """ElementVector: every component of the interpolated field is the
    scalar interpolation of that component's DOFs (w[c::dim])."""
    m = _omesh('tri', 6, 0.2)
    vb = O.CellBasis(m, O.ElementVector(O.ElementTriP2()))
    sb = O.CellBasis(m, O.ElementTriP2())
    w = np.random.default_rng(3).standard_normal(vb.N)
    val, grad = O.interpolate(vb, w)
    assert val.shape == (2,) + sb.dx.shape and grad.shape == (2, 2) + sb.dx.shape
    for c in range(2):
        sv, sg = O.interpolate(sb, w[c::2])
        assert np.array_equal(val[c], sv) and np.array_equal(grad[c], sg)


def test_asm_argument_checks():
    m = _tri(5)
    b0 = femhip.CellBasis(m, femhip.ElementTriP1(), elements=np.arange(4))
    b1 = femhip.CellBasis(m, femhip.ElementTriP1(), elements=np.arange(4, 8))
    from femhip.models import laplace
    with pytest.raises(NotImplementedError):
        femhip.asm(laplace, [b0, b1], [b1, b0])
    with pytest.raises(ValueError):
        femhip.asm(laplace, [])
    with pytest.raises(ValueError):
        femhip.asm(laplace)


@pytest.mark.parametrize('el', ['ElementTriP1', 'ElementTriP2'])
def test_vector_dofs_interleave_components(el):
    """Local function dim*k + c of ElementVector(e) is component c of scalar
    function k, and its global DOF is dim * (scalar DOF) + c: component c of
    a vector DOF array is w[c::dim] (the vector interpolation uses this)."""
    m = _tri(6)
    sb = femhip.CellBasis(m, getattr(femhip, el)())
    vb = femhip.CellBasis(m, femhip.ElementVector(getattr(femhip, el)()))
    assert vb.N == 2 * sb.N
    for k in range(sb.Nbfun):
        for c in range(2):
            assert np.array_equal(vb.element_dofs[2 * k + c], 2 * sb.element_dofs[k] + c)
    om = O.MeshTri(m.p, m.t)
    ovd, _ = O.element_dofs(om, O.ElementVector(getattr(O, el)()))
    assert np.array_equal(ovd, vb.element_dofs)


def test_default_mapping_is_accepted_others_refused():
    """``mapping=``: the mesh's default mapping (``Mesh.mapping()``,
    ``mesh.py:513-528``) is what the kernels evaluate, so it is accepted;
    any other mapping raises NotImplementedError (no CPU fallback)."""
    m = _tri(5)
    b = femhip.CellBasis(m, femhip.ElementTriP1(), mapping=m.mapping())
    assert isinstance(b.mapping, femhip.MappingAffine) and b.mapping.mesh is m
    femhip.CellBasis(m, femhip.ElementTriP2(), mapping=femhip.MappingAffine(m))
    twin = femhip.MeshTri(m.p.copy(), m.t.copy())
    femhip.CellBasis(m, femhip.ElementTriP1(), mapping=femhip.MappingAffine(twin))
    with pytest.raises(NotImplementedError):
        femhip.CellBasis(m, femhip.ElementTriP1(),
                         mapping=femhip.MappingIsoparametric(m, femhip.ElementTriP2()))
    other = _tri(6)
    with pytest.raises(NotImplementedError):
        femhip.CellBasis(m, femhip.ElementTriP1(), mapping=femhip.MappingAffine(other))
    x = np.linspace(0, 1, 3)
    h = femhip.MeshHex.init_tensor(x, x, x)
    assert isinstance(h.mapping(), femhip.MappingIsoparametric)
    femhip.CellBasis(h, femhip.ElementHex1(), mapping=h.mapping())
    with pytest.raises(NotImplementedError):
        femhip.CellBasis(h, femhip.ElementHex1(), mapping=femhip.MappingAffine(h))
    with pytest.raises(NotImplementedError):
        femhip.FacetBasis(m, femhip.ElementTriP1(), mapping=object())
