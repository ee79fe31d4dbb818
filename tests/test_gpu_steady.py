"""Later calls of scalar forms that have no fused kernel (P2 with
``w.c * u * v``, P1 with a per-point coefficient or ``w.c * u * v``, Q1/Q2
hexahedra): the validated pattern is reused without the exact-zero
signature pass.  The values and the summation order are the first call's, so
the CSR must be that call's bit for bit.  P2 laplace, mass and kappa-Laplace
with kappa per element later run the fused rows kernel (fixed-point sums):
the same pattern, values within 1e-12, bitwise repeatable."""
import numpy as np
import pytest

import femhip
from femhip import BilinearForm
from femhip.models import laplace, mass, kappa_laplace

pytestmark = pytest.mark.gpu

ELEM = {'tri1': femhip.ElementTriP1, 'tri2': femhip.ElementTriP2,
        'tet1': femhip.ElementTetP1, 'tet2': femhip.ElementTetP2,
        'hex1': femhip.ElementHex1, 'hex2': femhip.ElementHex2}


@BilinearForm
def robin(u, v, w):
    return w['alpha'] * u * v


def _basis(kind, elem, n):
    x = np.linspace(0, 1, n)
    M = {'tri': femhip.MeshTri, 'tet': femhip.MeshTet, 'hex': femhip.MeshHex}[kind]
    m = (M.init_tensor(x, x) if kind == 'tri' else M.init_tensor(x, x, x)).jittered(
        0.2, 1.0 / (n - 1))
    return femhip.Basis(m, ELEM[elem]())


@pytest.mark.parametrize('kind,elem,n', [('tri', 'tri2', 33), ('tet', 'tet2', 9),
                                         ('tri', 'tri1', 17), ('tet', 'tet1', 7),
                                         ('hex', 'hex1', 6), ('hex', 'hex2', 4)])
def test_second_call_bit_identical(kind, elem, n):
    print('case', kind, elem, n, flush=True)
    b = _basis(kind, elem, n)
    x = np.asarray(b.global_coordinates())
    E, Q = x.shape[1], x.shape[2]
    rng = np.random.default_rng(3)
    kap = rng.lognormal(0, 1, (E, 1) if elem.endswith('2') else (E, Q))
    alpha = 1.0 + x[0] ** 2 + 0.5 * x[1]
    calls = [(kappa_laplace, {'kappa': kap}), (robin, {'alpha': alpha})]
    if not elem.startswith(('tri1', 'tet1')):   # P1 laplace / mass: fused kernel
        calls += [(laplace, {}), (mass, {})]
    d = b.device()
    for form, kw in calls:
        A1 = form.assemble(b, **kw)
        A2 = form.assemble(b, **kw)
        A3 = form.assemble(b, **kw)
        rows = elem.endswith('2') and form in (laplace, mass, kappa_laplace) and kind != 'hex'
        for A in (A2, A3):
            assert np.array_equal(A.indptr, A1.indptr)
            assert np.array_equal(A.indices, A1.indices)
            if rows:
                assert d.last_verdict[0] == 1
                assert np.abs(A.data - A1.data).max() <= 1e-12 * np.abs(A1.data).max()
                assert np.array_equal(A.data, A2.data)
            else:
                assert np.array_equal(A.data, A1.data)
    assert d.validated
