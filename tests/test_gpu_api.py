"""GPU parity of element subsets (``CellBasis(elements=...)``), ``asm``
over lists of bases, distinct trial and test bases, and interpolation of
ElementVector fields, against the oracle: patterns bit-exact, values within
1e-12 of the largest entry (the subset's exact path, fused P1 kernel and P2
rows kernel alike), interpolated fields bit for bit."""
import numpy as np
import pytest

import femhip
from femhip.models import laplace, mass, unit_load
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu
TOL = 1e-12

_M = {'tri': (femhip.MeshTri, O.MeshTri), 'tet': (femhip.MeshTet, O.MeshTet),
      'hex': (femhip.MeshHex, O.MeshHex)}


def _mesh(kind, n, a, seed=0):
    x = np.linspace(0, 1, n)
    M = _M[kind][0]
    m = M.init_tensor(x, x) if kind == 'tri' else M.init_tensor(x, x, x)
    return m.jittered(a, 1.0 / (n - 1), seed=seed) if a else m


def _omesh(m):
    return _M[m.kind][1](m.p, m.t)


def _same(A, R):
    assert A.shape == R.shape
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert np.abs(A.data - R.data).max() <= TOL * np.abs(R.data).max()


def _left(m):
    """Elements whose centroid lies left of x = 0.45 (a ragged subset)."""
    return lambda x: x[0] < 0.45


@pytest.mark.parametrize('kind,el,n,a,form', [
    ('tri', 'ElementTriP1', 17, 0.2, 'laplace'), ('tet', 'ElementTetP1', 7, 0.2, 'laplace'),
    ('tet', 'ElementTetP1', 6, 0.2, 'mass'), ('tri', 'ElementTriP2', 9, 0.2, 'laplace'),
    ('tet', 'ElementTetP2', 5, 0.2, 'laplace'), ('tet', 'ElementTetP2', 5, 0.2, 'mass'),
    ('hex', 'ElementHex1', 5, 0.15, 'laplace'),
])
def test_subset_assembly_matches_the_oracle(kind, el, n, a, form):
    print('case', kind, el, n, form, flush=True)
    m = _mesh(kind, n, a)
    b = femhip.CellBasis(m, getattr(femhip, el)(), elements=_left(m))
    assert 0 < b.nelems < m.nelements
    ob = O.CellBasis(_omesh(m), getattr(O, el)(), elements=b.tind)
    f = {'laplace': laplace, 'mass': mass}[form]
    R = O.bilinear_assemble({'laplace': O.form_laplace, 'mass': O.form_mass}[form], ob)
    for call in range(3):      # exact path, then the fused path where there is one
        _same(f.assemble(b), R)
    d = b.device()
    if kind != 'hex':          # the last call ran the fused P1 / P2 rows kernel
        assert form in (d.fused_last if el.endswith('P1') else d.rows_last)
        assert d.last_verdict[0] == 1
    lv = unit_load.assemble(b)
    lr = O.linear_assemble(O.form_unit_load, ob)
    assert np.abs(lv - lr).max() <= TOL * np.abs(lr).max()
    assert np.all(lv[np.setdiff1d(np.arange(b.N), b.element_dofs)] == 0.0)


@pytest.mark.parametrize('kind,el,n', [('tri', 'ElementTriP1', 17),
                                       ('tet', 'ElementTetP2', 5)])
def test_asm_over_a_list_of_bases(kind, el, n):
    """asm(form, [b0, b1]) over an element partition: the reference's
    pattern (the COO concatenation of both), values of the full assembly
    to rounding; linear forms add up."""
    print('case', kind, el, n, flush=True)
    m = _mesh(kind, n, 0.2)
    c = m.p[:, m.t].mean(axis=1)
    sel = np.sin(7 * c[0]) + np.cos(5 * c[1]) > 0.3
    b0 = femhip.CellBasis(m, getattr(femhip, el)(), elements=sel)
    b1 = femhip.CellBasis(m, getattr(femhip, el)(), elements=~sel)
    full = femhip.CellBasis(m, getattr(femhip, el)())
    om = _omesh(m)
    for form, oform in ((laplace, O.form_laplace), (mass, O.form_mass)):
        for _ in range(2):
            A = femhip.asm(form, [b0, b1])
            _same(A, O.bilinear_assemble(oform, O.CellBasis(om, getattr(O, el)())))
            _same(A, form.assemble(full))
    v = femhip.asm(unit_load, [b0, b1])
    r = O.linear_assemble(O.form_unit_load, O.CellBasis(om, getattr(O, el)()))
    assert np.abs(v - r).max() <= TOL * np.abs(r).max()
    # a plain function of (u, v, w) is wrapped as the reference does
    B = femhip.asm(lambda u, v, w: u * v, [b0, b1])
    _same(B, O.bilinear_assemble(O.form_mass, O.CellBasis(om, getattr(O, el)())))


@pytest.mark.parametrize('kind,el,n,a', [
    ('tri', 'ElementTriP1', 9, 0.2), ('tri', 'ElementTriP2', 7, 0.2),
    ('tet', 'ElementTetP1', 5, 0.2), ('tet', 'ElementTetP2', 4, 0.2),
    ('hex', 'ElementHex1', 4, 0.15),
])
def test_vector_interpolation_bitexact(kind, el, n, a):
    """ElementVector fields: value (dim, E, Q), grad (dim, dim, E, Q), equal
    to the oracle's restatement of the reference bit for bit."""
    print('case', kind, el, n, flush=True)
    m = _mesh(kind, n, a)
    b = femhip.CellBasis(m, femhip.ElementVector(getattr(femhip, el)()))
    w = np.random.default_rng(11).standard_normal(b.N)
    f = b.interpolate(w)
    ob = O.CellBasis(_omesh(m), O.ElementVector(getattr(O, el)()))
    val, grad = O.interpolate(ob, w)
    assert f.value.shape == val.shape and f.grad.shape == grad.shape
    assert np.array_equal(f.value, val) and np.array_equal(f.grad, grad)


def test_subset_interpolation_bitexact():
    m = _mesh('tet', 5, 0.2)
    b = femhip.CellBasis(m, femhip.ElementTetP2(), elements=_left(m))
    w = np.random.default_rng(12).standard_normal(b.N)
    f = b.interpolate(w)
    val, grad = O.interpolate(O.CellBasis(_omesh(m), O.ElementTetP2(), elements=b.tind), w)
    assert f.value.shape == (b.nelems, b.W.shape[0])
    assert np.array_equal(f.value, val) and np.array_equal(f.grad, grad)


@pytest.mark.parametrize('kind,eu,ev,n,form', [
    ('tri', 'ElementTriP2', 'ElementTriP1', 9, 'laplace'),
    ('tri', 'ElementTriP1', 'ElementTriP2', 9, 'mass'),
    ('tet', 'ElementTetP2', 'ElementTetP1', 4, 'laplace'),
    ('tet', 'ElementTetP1', 'ElementTetP2', 4, 'coef_mass'),
    ('hex', 'ElementHex2', 'ElementHex1', 3, 'laplace'),
    ('tet', 'vP2', 'vP1', 3, 'elasticity'),
])
def test_distinct_trial_and_test_bases(kind, eu, ev, n, form):
    """BilinearForm.assemble(ubasis, vbasis): rows from the test basis,
    shape (vbasis.N, ubasis.N), against the oracle's mixed assembly; a
    second call reuses the pattern and gives the same result."""
    print('case', kind, eu, ev, n, form, flush=True)
    from femhip.models import linear_elasticity
    m = _mesh(kind, n, 0.15 if kind == 'hex' else 0.2)
    om = _omesh(m)

    def el(name, mod):
        if name.startswith('v'):
            return mod.ElementVector(getattr(mod, 'ElementTet' + name[1:])())
        return getattr(mod, name)()
    order = {'tri': 4, 'tet': 4, 'hex': 4}[kind]
    ub = femhip.CellBasis(m, el(eu, femhip), intorder=order)
    vb = femhip.CellBasis(m, el(ev, femhip), intorder=order)
    oub = O.CellBasis(om, el(eu, O), intorder=order)
    ovb = O.CellBasis(om, el(ev, O), intorder=order)
    kw, okw = {}, {}
    if form == 'laplace':
        F, OF = laplace, O.form_laplace
    elif form == 'mass':
        F, OF = mass, O.form_mass
    elif form == 'coef_mass':
        c = 1.0 + np.random.default_rng(2).random((ub.nelems, ub.W.shape[0]))
        F = femhip.BilinearForm(lambda u, v, w: w.c * u * v)
        OF, kw, okw = O.form_coef_mass, {'c': c}, {'c': c}
    else:
        lam, mu = O.lame_parameters(1.0, 0.3)
        F, OF = linear_elasticity(lam, mu), O.form_elasticity(lam, mu)
    R = O.bilinear_assemble(OF, oub, vbasis=ovb, **okw)
    A = F.assemble(ub, vb, **kw)
    assert A.shape == (vb.N, ub.N) == R.shape
    _same(A, R)
    A2 = F.assemble(ub, vb, **kw)
    assert np.array_equal(A2.data, A.data) and np.array_equal(A2.indices, A.indices)


def test_asm_over_trial_and_test_lists():
    """asm(form, [u0, u1], [v0, v1]) sums every (trial, test) combination
    (assembly/__init__.py:91-94); here the two halves of an element
    partition with a P2 trial and a P1 test space on each."""
    m = _mesh('tri', 9, 0.2)
    c = m.p[:, m.t].mean(axis=1)
    sel = c[0] < 0.5
    ub = [femhip.CellBasis(m, femhip.ElementTriP2(), elements=s) for s in (sel, ~sel)]
    vb = [femhip.CellBasis(m, femhip.ElementTriP1(), intorder=4, elements=s)
          for s in (sel, ~sel)]
    with pytest.raises(NotImplementedError):   # (u0, v1): different elements
        femhip.asm(mass, ub, vb)
    A = femhip.asm(mass, ub[0], vb[0])
    om = _omesh(m)
    R = O.bilinear_assemble(O.form_mass, O.CellBasis(om, O.ElementTriP2(), elements=ub[0].tind),
                            vbasis=O.CellBasis(om, O.ElementTriP1(), intorder=4,
                                               elements=vb[0].tind))
    _same(A, R)


def test_default_mapping_argument_changes_nothing():
    """``CellBasis(mapping=mesh.mapping())`` assembles exactly what the
    default basis does (the kernels evaluate that mapping)."""
    for kind, el, n in (('tet', 'ElementTetP1', 5), ('hex', 'ElementHex1', 4)):
        m = _mesh(kind, n, 0.15)
        A = laplace.assemble(femhip.CellBasis(m, getattr(femhip, el)()))
        B = laplace.assemble(femhip.CellBasis(m, getattr(femhip, el)(), mapping=m.mapping()))
        assert np.array_equal(A.indices, B.indices) and np.array_equal(A.data, B.data)


@pytest.mark.parametrize('kind,el', [('tri', 'ElementTriP1'), ('tet', 'ElementTetP2')])
def test_empty_element_subset(kind, el):
    """No element selected: the all-zero matrix and load of the full DOF
    count, as the reference's assembly over zero elements gives."""
    m = _mesh(kind, 5, 0.2)
    b = femhip.CellBasis(m, getattr(femhip, el)(), elements=np.zeros(m.nelements, bool))
    assert b.nelems == 0
    for _ in range(2):
        A = laplace.assemble(b)
        assert A.shape == (b.N, b.N) and A.nnz == 0
    v = unit_load.assemble(b)
    assert v.shape == (b.N,) and not v.any()
    R = O.bilinear_assemble(O.form_laplace, O.CellBasis(_omesh(m), getattr(O, el)(),
                                                        elements=b.tind))
    assert R.nnz == 0 and R.shape == A.shape


def test_elasticity_on_an_irregular_element_subset():
    """Vector Q2-hex elasticity on an irregular element subset: CSR rows
    that hold 1 to 8 local rows (3, 5 and 6 occur only off a full tensor
    mesh), through the exact first call and then the MFMA local matrices
    with the row sums; pattern bit-exact, values within 1e-12 of the
    largest entry of the oracle's assembly on every call."""
    from femhip.models import linear_elasticity
    m = _mesh('hex', 5, 0.15)
    c = m.p[:, m.t].mean(axis=1)
    sel = np.sin(7 * c[0]) + np.cos(5 * c[1]) + np.sin(3 * c[2]) > 0.3
    b = femhip.CellBasis(m, femhip.ElementVector(femhip.ElementHex2()), intorder=5,
                         elements=sel)
    assert 0 < b.nelems < m.nelements and b.W.shape[0] == 27
    lam, mu = O.lame_parameters(1.0, 0.3)
    R = O.bilinear_assemble(O.form_elasticity(lam, mu),
                            O.CellBasis(_omesh(m), O.ElementVector(O.ElementHex2()),
                                        intorder=5, elements=b.tind))
    F = linear_elasticity(lam, mu)
    for call in range(3):
        print('call', call, flush=True)
        _same(F.assemble(b), R)
    d = b.device()
    plan = next(iter(d.validated.values()))
    rs = getattr(plan, 'row_sums', None)
    assert rs is not None                      # the row sums ran
    counts = set(np.diff(rs[1].cpu().numpy()).tolist())
    assert {3, 5, 6} <= counts
