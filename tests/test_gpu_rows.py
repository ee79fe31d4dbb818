"""GPU parity of the fused rows kernel (P2 triangles and tetrahedra,
laplace, mass and kappa-Laplace): later calls run it on the validated
pattern and must give the oracle's pattern and values within 1e-12, also
for new coordinates and on unjittered meshes whose entries include
structural zeros (every element's near-zero entries rechecked), and fall
back to the exact path when the new input's exact zeros differ from the
pattern's."""
import numpy as np
import pytest
import torch

import femhip
from femhip.models import laplace, mass
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _mesh(kind, n, a, seed=0):
    x = np.linspace(0, 1, n)
    m = (femhip.MeshTri.init_tensor(x, x) if kind == 'tri'
         else femhip.MeshTet.init_tensor(x, x, x))
    return m.jittered(a, 1.0 / (n - 1), seed=seed) if a else m


def _oracle(kind, p, t, form):
    om = (O.MeshTri if kind == 'tri' else O.MeshTet)(p, t)
    ob = O.CellBasis(om, O.ElementTriP2() if kind == 'tri' else O.ElementTetP2())
    return O.bilinear_assemble({'laplace': O.form_laplace, 'mass': O.form_mass}[form], ob)


def _same(A, R):
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert np.abs(A.data - R.data).max() <= TOL * np.abs(R.data).max()


@pytest.mark.parametrize('kind,n,a,form', [
    ('tri', 17, 0.2, 'laplace'), ('tri', 17, 0.0, 'laplace'), ('tri', 9, 0.2, 'mass'),
    ('tri', 9, 0.0, 'mass'), ('tet', 6, 0.2, 'laplace'), ('tet', 5, 0.0, 'laplace'),
    ('tet', 5, 0.2, 'mass'), ('tet', 5, 0.0, 'mass'),
])
def test_rows_kernel_matches_the_oracle(kind, n, a, form):
    print('case', kind, n, a, form, flush=True)
    m = _mesh(kind, n, a)
    b = femhip.Basis(m, femhip.ElementTriP2() if kind == 'tri' else femhip.ElementTetP2())
    f = {'laplace': laplace, 'mass': mass}[form]
    R = _oracle(kind, m.p, m.t, form)
    _same(f.assemble(b), R)                    # exact path, pattern validated
    d = b.device()
    A2 = f.assemble(b)                         # fused rows kernel
    assert form in d.rows_last and d.last_verdict[0] == 1
    _same(A2, R)
    A3 = f.assemble(b)
    assert np.array_equal(A3.data, A2.data)    # bitwise repeatable


@pytest.mark.parametrize('kind,n', [('tri', 17), ('tet', 6)])
def test_rows_kernel_new_coordinates(kind, n):
    """New interior coordinates: the same pattern (no new exact zeros), the
    fused pass validates it and gives the oracle's values for them."""
    m = _mesh(kind, n, 0.2)
    b = femhip.Basis(m, femhip.ElementTriP2() if kind == 'tri' else femhip.ElementTetP2())
    laplace.assemble(b)
    laplace.assemble(b)
    d = b.device()
    for seed in (5, 6):
        m2 = _mesh(kind, n, 0.2, seed=seed)
        print('coordinates seed', seed, flush=True)
        d.set_coordinates(torch.tensor(m2.p, device='cuda'))
        A = laplace.assemble(b)
        assert d.last_verdict[0] == 1
        _same(A, _oracle(kind, m2.p, m2.t, 'laplace'))


def test_rows_kernel_falls_back_when_the_zeros_change():
    """The unjittered mesh prunes exact zeros; jittered coordinates make
    those entries nonzero: the fused pass flags the input (a nonzero entry
    outside the pattern) and the call returns the exact path's result."""
    m = _mesh('tet', 5, 0.0)
    b = femhip.Basis(m, femhip.ElementTetP2())
    R0 = _oracle('tet', m.p, m.t, 'laplace')
    _same(laplace.assemble(b), R0)
    _same(laplace.assemble(b), R0)
    d = b.device()
    assert d.last_verdict[0] == 1
    m2 = _mesh('tet', 5, 0.2, seed=3)
    d.set_coordinates(torch.tensor(m2.p, device='cuda'))
    R1 = _oracle('tet', m2.p, m2.t, 'laplace')
    assert R1.nnz != R0.nnz
    A = laplace.assemble(b)
    assert d.last_verdict[0] == 0
    _same(A, R1)
    # the new zero set is kept: the calls that follow on these coordinates
    # get a rows plan of the new pattern and run fused again
    for _ in range(2):
        _same(laplace.assemble(b), R1)
    assert d.last_verdict[0] == 1 and d.rows_last['laplace'].nnz == R1.nnz


@pytest.mark.parametrize('kind,n', [('tri', 17), ('tet', 6)])
def test_rows_kernel_kappa_laplace(kind, n):
    """kappa-Laplace with kappa per element: later calls run the rows
    kernel (the metric times kappa_e); a new kappa that keeps every exact
    zero (twice the first: every rounding scales exactly) stays on it, an
    unrelated new kappa may change the reference's zeros on the mesh's
    right-angled boundary elements (either verdict, the result is the
    reference's), and a kappa with zeros falls back to the exact path;
    every result against the oracle."""
    from femhip.models import kappa_laplace
    m = _mesh(kind, n, 0.2)
    b = femhip.Basis(m, femhip.ElementTriP2() if kind == 'tri' else femhip.ElementTetP2())
    om = (O.MeshTri if kind == 'tri' else O.MeshTet)(m.p, m.t)
    ob = O.CellBasis(om, O.ElementTriP2() if kind == 'tri' else O.ElementTetP2())
    rng = np.random.default_rng(4)
    E = m.nelements
    k1 = rng.lognormal(0, 1, (E, 1))
    k2 = rng.lognormal(0, 1, (E, 1))
    k3 = k2.copy()
    k3[rng.random(E) < 0.1] = 0.0
    d = b.device()
    steps = ((k1, False), (k1, True), (2.0 * k1, True), (k2, None), (k3, False))
    for step, (k, fused) in enumerate(steps):
        print('step', step, flush=True)
        A = kappa_laplace.assemble(b, kappa=k)
        R = O.bilinear_assemble(O.form_kappa_laplace, ob, kappa=k)
        _same(A, R)
        if step > 0:
            assert 'kappa_laplace' in d.rows_last
            if fused is not None:
                assert (d.last_verdict[0] == 1) == fused, d.last_verdict
