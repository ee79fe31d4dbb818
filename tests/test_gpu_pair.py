"""Laplace, mass and unit load of a P1 basis in one fused pass: the plan is
built on the mass pattern and the Laplace values are written compacted to
the Laplace pattern (which lacks the slots whose contributions are all
exactly zero, on unjittered meshes); each output against the oracle."""
import numpy as np
import pytest
import torch

import femhip
from femhip.fused import LaplaceMassPass
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check(ip, ix, data, R):
    assert np.array_equal(ip.cpu().numpy(), R.indptr)
    assert np.array_equal(ix.cpu().numpy(), R.indices)
    assert np.abs(data.cpu().numpy() - R.data).max() <= TOL * np.abs(R.data).max()


@pytest.mark.parametrize('kind,n,a', [('tet', 9, 0.2), ('tet', 8, 0.0), ('tri', 33, 0.2),
                                      ('tri', 17, 0.0)])
def test_laplace_mass_load_in_one_pass(kind, n, a):
    print('case', kind, n, a, flush=True)
    x = np.linspace(0, 1, n)
    m = (femhip.MeshTri.init_tensor(x, x) if kind == 'tri'
         else femhip.MeshTet.init_tensor(x, x, x))
    if a:
        m = m.jittered(a, 1.0 / (n - 1))
    el = femhip.ElementTriP1() if kind == 'tri' else femhip.ElementTetP1()
    b = femhip.Basis(m, el)
    pair = LaplaceMassPass(b)
    L, M, f = pair.assemble()
    torch.cuda.synchronize()
    assert pair.fp.verdict.cpu().tolist()[0] == 1
    om = (O.MeshTri if kind == 'tri' else O.MeshTet)(m.p, m.t)
    ob = O.CellBasis(om, O.ElementTriP1() if kind == 'tri' else O.ElementTetP1())
    RL = O.bilinear_assemble(O.form_laplace, ob)
    RM = O.bilinear_assemble(O.form_mass, ob)
    rf = O.linear_assemble(O.form_unit_load, ob)
    if not a and kind == 'tet':
        assert RL.nnz < RM.nnz              # the compaction is exercised
    _check(pair.plan_l.indptr, pair.plan_l.indices, L, RL)
    _check(pair.plan_m.indptr, pair.plan_m.indices, M, RM)
    assert np.abs(f.cpu().numpy() - rf).max() <= TOL * np.abs(rf).max()
    L2, M2, f2 = pair.assemble()
    assert torch.equal(L, L2) and torch.equal(M, M2) and torch.equal(f, f2)
