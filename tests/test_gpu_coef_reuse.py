"""Coefficient reuse across calls (plan validation must follow the
coefficient's contents, not its identity):

* a loop over freshly allocated device kappa tensors -- the caching
  allocator hands out the same address, and a fresh tensor's version
  counter is 0 -- on P1 (fused path) and P2 (exact path) meshes, one of the
  kappas holding exact zeros (which prune entries) -- each result against
  the oracle (pattern bit-exact, values within 1e-12 of max|A|);
* one DOF vector used as the coefficient on two facet bases of equal size
  (the left and right sides of the unit square) -- each against the oracle.
"""
import numpy as np
import pytest
import torch

import femhip
from femhip import LinearForm
from femhip.models import kappa_laplace
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu


def _close(a, ref, tol=1e-12):
    return np.abs(a - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize('elem,n', [('P1', 9), ('P2', 7)])
def test_fresh_device_kappa_every_call(elem, n):
    x = np.linspace(0, 1, n)
    m = femhip.MeshTet.init_tensor(x, x, x)
    E = m.nelements
    fe = {'P1': femhip.ElementTetP1, 'P2': femhip.ElementTetP2}[elem]
    b = femhip.Basis(m, fe())
    ob = O.CellBasis(O.MeshTet(m.p, m.t), getattr(O, fe.__name__)())
    rng = np.random.default_rng(11)
    kappas = [rng.lognormal(0, 1, E), rng.lognormal(0, 1, E), rng.lognormal(0, 1, E)]
    kappas[1][::3] = 0.0           # exact zeros prune entries
    kappas.append(kappas[0].copy())
    for step, k in enumerate(kappas):
        print('case', elem, 'step', step, flush=True)
        kt = torch.tensor(k, device='cuda')          # fresh allocation each call
        A = kappa_laplace.assemble(b, kappa=kt[:, None])
        R = O.bilinear_assemble(O.form_kappa_laplace, ob,
                                kappa=np.broadcast_to(k[:, None], ob.dx.shape))
        assert np.array_equal(A.indptr, R.indptr), step
        assert np.array_equal(A.indices, R.indices), step
        assert _close(A.data, R.data), step
        del kt, A


@LinearForm
def neumann(v, w):
    return w['g'] * v


def test_dof_coefficient_on_two_equal_facet_sets():
    x = np.linspace(0, 1, 9)
    m = femhip.MeshTri.init_tensor(x, x).jittered(0.2, 1 / 8)
    om = O.MeshTri(m.p, m.t)
    left = np.nonzero(np.all(np.isclose(m.p[0, m.facets], 0.0), axis=0))[0]
    right = np.nonzero(np.all(np.isclose(m.p[0, m.facets], 1.0), axis=0))[0]
    assert len(left) == len(right) > 0
    w = np.random.default_rng(2).standard_normal(m.p.shape[1])
    for name, f in (('left', left), ('right', right)):
        print('case', name, flush=True)
        fb = femhip.FacetBasis(m, femhip.ElementTriP1(), facets=f)
        ofb = O.FacetBasis(om, O.ElementTriP1(), facets=f)
        ref = O.linear_assemble(O.form_coef_load, ofb, c=O.interpolate(ofb, w)[0])
        got = neumann.assemble(fb, g=w)
        assert _close(got, ref), name
