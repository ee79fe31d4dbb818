"""The reference-side ctypes module of the integration guide (the first
Python block of the guide at the repository root), executed as written
(only the library path is pointed at this build): the P1 tetrahedral
Laplace and P2 triangular mass matrices it returns equal the oracle's."""
import os
import re

import numpy as np
import pytest

import femhip
from femhip import _lib
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stub():
    text = open(os.path.join(ROOT, 'INTEGRATION.md')).read()
    m = re.search(r'<!-- integration-stub -->\s*```python\n(.*?)```\s*<!-- /integration-stub -->',
                  text, re.S)
    assert m, "stub block not found"
    code = m.group(1).replace('ctypes.CDLL("libfemhip.so")',
                              'ctypes.CDLL(%r)' % _lib.LIB_PATH)
    ns = {}
    exec(compile(code, 'INTEGRATION.md', 'exec'), ns)
    return ns


@pytest.mark.parametrize('kind', ['tet1', 'tri2'])
def test_integration_stub_runs_and_matches(kind):
    ns = _stub()
    if kind == 'tet1':
        x = np.linspace(0, 1, 6)
        m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 0.2)
        e, oe, om, form = femhip.ElementTetP1(), O.ElementTetP1(), O.MeshTet, 'laplace'
    else:
        x = np.linspace(0, 1, 9)
        m = femhip.MeshTri.init_tensor(x, x)
        e, oe, om, form = femhip.ElementTriP2(), O.ElementTriP2(), O.MeshTri, 'mass'
    b = femhip.Basis(m, e)
    phi, dphi, _ = b.tables()
    print('stub', kind, flush=True)
    A = ns['assemble_builtin'](ns['LAPLACE'] if form == 'laplace' else ns['MASS'], m.p, m.t,
                               b.element_dofs, b.N, (b.X, b.W), phi, dphi)
    R = O.assemble_named(form, O.CellBasis(om(m.p, m.t), oe))
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert np.abs(A.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
