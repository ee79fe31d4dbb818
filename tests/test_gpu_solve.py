"""The assembled system kept on the device for the solve: condense (rows and
columns of the Dirichlet DOFs removed, right-hand side corrected), y = A x
and the Jacobi-preconditioned CG, against the oracle's condense and SciPy."""
import numpy as np
import pytest
import torch
from scipy.sparse.linalg import spsolve

import femhip
from femhip.models import laplace, unit_load
from femhip.utils import condense_device, solve_device, spmv
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu


def _tri(n):
    x = np.linspace(0, 1, n)
    return femhip.MeshTri.init_tensor(x, x)


def test_condense_matches_oracle():
    m = _tri(33)
    b = femhip.Basis(m, femhip.ElementTriP1())
    A = laplace.assemble_device(b)
    f = unit_load.assemble(b)
    D = m.boundary_nodes()
    xb = np.zeros(b.N)
    xb[D] = np.sin(3 * m.p[0, D]) + m.p[1, D]
    print('condense', flush=True)
    AII, bI, xx, I = condense_device(A, f, x=xb, D=D)
    torch.cuda.synchronize()
    Ar = A.to_scipy()
    RII, rI, _, RI = O.condense(Ar, f, x=xb, D=D)
    assert np.array_equal(I.cpu().numpy(), RI)
    assert np.array_equal(AII.indptr.cpu().numpy(), RII.indptr)
    assert np.array_equal(AII.indices.cpu().numpy(), RII.indices)
    assert np.array_equal(AII.data.cpu().numpy(), RII.data)
    assert np.abs(bI.cpu().numpy() - rI).max() <= 1e-14 * np.abs(rI).max()
    # I given instead of D
    AII2, bI2, _, I2 = condense_device(A, f, x=xb, I=RI)
    assert torch.equal(AII2.indices, AII.indices) and torch.equal(bI2, bI)


def test_spmv_matches_scipy():
    x = np.linspace(0, 1, 9)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1 / 8)
    b = femhip.Basis(m, femhip.ElementTetP2())
    A = laplace.assemble_device(b)
    v = np.random.default_rng(4).standard_normal(b.N)
    y = spmv(A, torch.tensor(v, device='cuda')).cpu().numpy()
    ref = A.to_scipy() @ v
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_solve_ex01_on_device():
    """ex01 (Poisson, unit load, zero boundary values) solved on the device:
    max(u) = 0.073657185490792 (the reference's test_examples value) and the
    SciPy direct solution of the same condensed system."""
    m = _tri(65)
    b = femhip.Basis(m, femhip.ElementTriP1())
    A = laplace.assemble_device(b)
    f = unit_load.assemble(b)
    AII, bI, xx, I = condense_device(A, f, I=m.interior_nodes())
    info = {}
    print('pcg', flush=True)
    u = solve_device(AII, bI, xx, I, rtol=1e-13, info=info)
    u2 = solve_device(AII, bI, xx, I, rtol=1e-13)
    torch.cuda.synchronize()
    assert torch.equal(u, u2), "the CG iterates are not bitwise repeatable"
    un = u.cpu().numpy()
    assert abs(un.max() - 0.073657185490792) < 1e-9
    Ar = AII.to_scipy()
    ud = spsolve(Ar.tocsc(), bI.cpu().numpy())
    assert np.abs(un[m.interior_nodes()] - ud).max() <= 1e-10 * np.abs(ud).max()
    assert 0 < info['iterations'] < 1000


def test_solve_tet_residual():
    """3-D Laplace on jittered P1 tetrahedra with nonzero boundary values:
    the residual of the device solution, checked with SciPy on the host."""
    x = np.linspace(0, 1, 25)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1 / 24)
    b = femhip.Basis(m, femhip.ElementTetP1())
    A = laplace.assemble_device(b)
    f = unit_load.assemble(b)
    D = m.boundary_nodes()
    xb = np.zeros(b.N)
    xb[D] = m.p[0, D] * m.p[1, D]
    AII, bI, xx, I = condense_device(A, f, x=xb, D=D)
    info = {}
    u = solve_device(AII, bI, xx, I, rtol=1e-11, info=info)
    Ar = A.to_scipy().tocsr()
    un = u.cpu().numpy()
    Ii = I.cpu().numpy()
    r = (Ar @ un - f)[Ii]
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(bI.cpu().numpy())
    assert np.array_equal(un[D], xb[D])


def test_pcg_diagonal_system_stays_finite():
    """Jacobi-PCG on a diagonal system converges in one step to r = 0
    exactly; the following iterations (up to the next residual check) must
    leave x unchanged instead of forming 0/0.  Also the 1-DOF system."""
    from femhip.device import DeviceCSR
    for diag in ([2.0, 4.0, 8.0], [5.0]):
        n = len(diag)
        dev = 'cuda'
        A = DeviceCSR(torch.arange(n + 1, dtype=torch.int32, device=dev),
                      torch.arange(n, dtype=torch.int32, device=dev),
                      torch.tensor(diag, dtype=torch.float64, device=dev), (n, n))
        b = torch.tensor([1.0, -3.0, 0.5][:n], dtype=torch.float64, device=dev)
        info = {}
        print('pcg diag', n, flush=True)
        u = solve_device(A, b, rtol=1e-12, check_every=16, info=info).cpu().numpy()
        assert np.all(np.isfinite(u)), u
        assert np.allclose(u, b.cpu().numpy() / np.array(diag), rtol=0, atol=1e-15)
        assert info['residual'] == 0.0
