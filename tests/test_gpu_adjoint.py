"""Adjoint of the P1 kappa-Laplace + load assembly against FP64 central
differences, the closed form dL/dkappa_e = 1/2 u_e^T K_e(1) u_e, and the
reference gradient written from scikit-fem itself (65^3 mesh)."""
import os

import numpy as np
import pytest
import torch

import femhip
from femhip.autograd import kappa_laplace_with_load, quadratic_loss
from oracle import skfem_np as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
pytestmark = pytest.mark.gpu


def _setup(kind, n, a, seed=0):
    x = np.linspace(0, 1, n)
    if kind == 'tri':
        m = femhip.MeshTri.init_tensor(x, x)
        el = femhip.ElementTriP1()
    else:
        m = femhip.MeshTet.init_tensor(x, x, x)
        el = femhip.ElementTetP1()
    if a:
        m = m.jittered(a, 1.0 / (n - 1))
    b = femhip.Basis(m, el)
    rng = np.random.default_rng(seed)
    dev = torch.device('cuda')
    kappa = torch.tensor(rng.lognormal(0, 1, m.nelements), device=dev)
    u = torch.tensor(rng.standard_normal(b.N), device=dev)
    f = torch.tensor(rng.standard_normal(b.N), device=dev)
    g = torch.tensor(rng.standard_normal(b.N), device=dev)
    p = torch.tensor(m.p, device=dev)
    return m, b, kappa, u, f, g, p


def _loss(b, kappa, p, u, f, g):
    data, load, (ip, ix) = kappa_laplace_with_load(b, kappa, p)
    return quadratic_loss(data, ip, ix, u, f) + (g * load).sum()


@pytest.mark.parametrize('kind,n,a', [('tri', 9, 0.2), ('tet', 5, 0.2), ('tet', 4, 0.0)])
def test_adjoint_vs_central_differences(kind, n, a):
    m, b, kappa, u, f, g, p = _setup(kind, n, a)
    kappa.requires_grad_(True)
    p.requires_grad_(True)
    L = _loss(b, kappa, p, u, f, g)
    L.backward()
    dk, dX = kappa.grad.clone(), p.grad.clone()
    h = 1e-6
    with torch.no_grad():
        fd_k = torch.empty_like(dk)
        for e in range(m.nelements):
            kp, km = kappa.detach().clone(), kappa.detach().clone()
            kp[e] += h
            km[e] -= h
            fd_k[e] = (_loss(b, kp, p.detach(), u, f, g) - _loss(b, km, p.detach(), u, f, g)) / (2 * h)
        assert (dk - fd_k).abs().max() <= 1e-7 * dk.abs().max()
        rng = np.random.default_rng(1)
        nodes = rng.choice(m.p.shape[1], size=min(24, m.p.shape[1]), replace=False)
        for nd in nodes:
            for c in range(m.dim):
                pp, pm = p.detach().clone(), p.detach().clone()
                pp[c, nd] += h
                pm[c, nd] -= h
                fd = (_loss(b, kappa.detach(), pp, u, f, g)
                      - _loss(b, kappa.detach(), pm, u, f, g)) / (2 * h)
                assert abs(float(dX[c, nd] - fd)) <= 1e-7 * float(dX.abs().max()), (nd, c)


def test_dkappa_closed_form():
    m, b, kappa, u, f, g, p = _setup('tet', 6, 0.2, seed=3)
    kappa.requires_grad_(True)
    data, load, (ip, ix) = kappa_laplace_with_load(b, kappa, p)
    quadratic_loss(data, ip, ix, u, f).backward()
    ob = O.CellBasis(O.MeshTet(m.p, m.t), O.ElementTetP1())
    _, K, _ = O.bilinear_coo(O.form_laplace, ob)
    K = K.reshape(4, 4, -1)
    uu = u.cpu().numpy()[m.t]  # (4, E)
    ref = 0.5 * np.einsum('ie,jie,je->e', uu, K, uu)
    got = kappa.grad.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_reference_gradient_65():
    """Config: kappa-Laplace on MeshTet 65^3, kappa ~ lognormal(0,1) seed 0,
    u, f ~ N(0,1) seeds 1, 2; reference dL/dkappa (single-element skfem
    assemblies) and dL/dX (skfem central differences, step 1e-6)."""
    z = np.load(os.path.join(GOLD, 'grad_kappa65.npz'))
    n = int(z['n'])
    x = np.linspace(0, 1, n)
    m = femhip.MeshTet.init_tensor(x, x, x)
    b = femhip.Basis(m, femhip.ElementTetP1())
    dev = torch.device('cuda')
    kappa = torch.tensor(np.random.default_rng(0).lognormal(0.0, 1.0, m.nelements),
                         device=dev, requires_grad=True)
    u = torch.tensor(np.random.default_rng(1).standard_normal(b.N), device=dev)
    f = torch.tensor(np.random.default_rng(2).standard_normal(b.N), device=dev)
    p = torch.tensor(m.p, device=dev, requires_grad=True)
    data, load, (ip, ix) = kappa_laplace_with_load(b, kappa, p)
    L = quadratic_loss(data, ip, ix, u, f)
    assert abs(float(L) - float(z['L'])) <= 1e-12 * abs(float(z['L']))
    L.backward()
    dk = kappa.grad.cpu().numpy()[z['elements']]
    assert np.abs(dk - z['dkappa']).max() <= 1e-10 * np.abs(z['dkappa']).max()
    dX = p.grad.cpu().numpy()[:, z['nodes']].T
    ref = z['dX']
    assert np.abs(dX - ref).max() <= 1e-7 * np.abs(ref).max()


@pytest.mark.parametrize('kind,n', [('tri', 17), ('tet', 7)])
def test_adjoint_is_bitwise_repeatable(kind, n):
    """The node sums are fixed point with one exponent per launch: two
    backward passes on the same upstream give identical gradients, also
    after a pass on another upstream changed the exponent."""
    m, b, kappa, u, f, g, p = _setup(kind, n, 0.2, seed=7)
    kappa.requires_grad_(True)
    p.requires_grad_(True)
    data, load, _ = kappa_laplace_with_load(b, kappa, p)
    gd = torch.randn_like(data)
    gl = torch.randn_like(load)
    r1 = torch.autograd.grad((data, load), (kappa, p), (gd, gl), retain_graph=True)
    torch.autograd.grad((data, load), (kappa, p), (1e6 * gd, gl), retain_graph=True)
    r2 = torch.autograd.grad((data, load), (kappa, p), (gd, gl), retain_graph=True)
    assert torch.equal(r1[0], r2[0]) and torch.equal(r1[1], r2[1])
    # linearity in the upstream (to the fixed-point rounding)
    r3 = torch.autograd.grad((data, load), (kappa, p), (2 * gd, 2 * gl), retain_graph=True)
    assert (r3[1] - 2 * r1[1]).abs().max() <= 1e-12 * r1[1].abs().max()
    assert (r3[0] - 2 * r1[0]).abs().max() <= 1e-13 * r1[0].abs().max()
