"""The oracle pinned against the reference's golden data and against the
mathematics (CPU only)."""
import glob
import json
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from oracle import skfem_np as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
EL = {'tri1': O.ElementTriP1, 'tri2': O.ElementTriP2, 'tet1': O.ElementTetP1,
      'tet2': O.ElementTetP2, 'hex1': O.ElementHex1, 'hex2': O.ElementHex2}
MESH = {'tri': O.MeshTri, 'tet': O.MeshTet, 'hex': O.MeshHex}


def _fixtures():
    return sorted(glob.glob(os.path.join(GOLD, 'small_*.npz')))


@pytest.mark.parametrize('path', _fixtures(), ids=lambda p: os.path.basename(p)[6:-4])
def test_oracle_matches_reference_fixture(path):
    z = np.load(path)
    meta = json.loads(str(z['meta']))
    name = os.path.basename(path)[6:-4]
    m = MESH[meta['kind']](z['p'], z['t'])
    e = EL[meta['element']]()
    if meta['vector']:
        e = O.ElementVector(e)
    b = O.CellBasis(m, e, intorder=None if meta['intorder'] < 0 else meta['intorder'])
    assert np.array_equal(b.element_dofs, z['element_dofs'])
    assert b.N == int(z['N'])
    lam, mu = O.lame_parameters(1.0, 0.3)
    for form in meta['forms']:
        kw = {}
        if form == 'kappa_laplace':
            kw['kappa'] = np.repeat(z['kappa_e'][:, None], b.X.shape[1], axis=1)
        if form == 'linear_elasticity':
            kw = dict(Lambda=lam, Mu=mu)
        r = O.assemble_named(form, b, **kw)
        if form == 'unit_load':
            ref = z['unit_load_b']
            assert np.array_equal(r, ref)
            continue
        assert np.array_equal(r.indptr, z[form + '_indptr'])
        assert np.array_equal(r.indices, z[form + '_indices'])
        assert np.array_equal(r.data, z[form + '_data'])


def test_oracle_config1():
    z = np.load(os.path.join(GOLD, 'config1_tri129.npz'))
    x = np.linspace(0, 1, 129)
    m = O.MeshTri.init_tensor(x, x)
    b = O.CellBasis(m, O.ElementTriP1())
    A = O.bilinear_assemble(O.form_laplace, b)
    f = O.linear_assemble(O.form_unit_load, b)
    assert A.nnz == 82689
    assert np.array_equal(A.indptr, z['indptr']) and np.array_equal(A.indices, z['indices'])
    assert np.array_equal(A.data, z['data'])
    assert np.array_equal(f, z['b'])
    assert np.array_equal(m.boundary_nodes(), z['D'])


def test_doctest_mass_matrix():
    """skfem/assembly/__init__.py:38-46 and bilinear_form.py:24-30."""
    b = O.CellBasis(O.MeshTri(), O.ElementTriP1())
    M = O.bilinear_assemble(O.form_mass, b).toarray()
    ref = np.array([[0.08333333, 0.04166667, 0.04166667, 0.],
                    [0.04166667, 0.16666667, 0.08333333, 0.04166667],
                    [0.04166667, 0.08333333, 0.16666667, 0.04166667],
                    [0., 0.04166667, 0.04166667, 0.08333333]])
    assert np.abs(M - ref).max() < 5e-9
    # l(v) = x^2 v from the same doctest
    xq = b.mapping.F(b.X)[0]
    f = O.linear_assemble(lambda v, w: xq ** 2 * v[0], b)
    assert np.abs(f - np.array([0.0162037, 0.15046296, 0.06712963,
                                0.09953704])).max() < 5e-9


def test_ex01_max_oracle():
    from scipy.sparse.linalg import spsolve
    x = np.linspace(0, 1, 65)
    m = O.MeshTri.init_tensor(x, x)
    b = O.CellBasis(m, O.ElementTriP1())
    A = O.bilinear_assemble(O.form_laplace, b)
    f = O.linear_assemble(O.form_unit_load, b)
    A, f = O.enforce(A, f, D=m.boundary_nodes())
    u = spsolve(A.tocsc(), f)
    assert abs(u.max() - 0.073657185490792) < 1e-7


_DOFLOCS = {
    'ElementTriP1': [(0, 0), (1, 0), (0, 1)],
    'ElementTriP2': [(0, 0), (1, 0), (0, 1), (.5, 0), (.5, .5), (0, .5)],
    'ElementTetP1': [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    'ElementTetP2': [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (.5, 0, 0), (.5, .5, 0),
                     (0, .5, 0), (0, 0, .5), (.5, 0, .5), (0, .5, .5)],
    'ElementHex1': [tuple(r) for r in O.HEX2_DOFLOCS[:8]],
    'ElementHex2': [tuple(r) for r in O.HEX2_DOFLOCS],
}


@pytest.mark.parametrize('kind,elem,deg', [
    ('tri', 'ElementTriP1', 1), ('tri', 'ElementTriP2', 2), ('tet', 'ElementTetP1', 1),
    ('tet', 'ElementTetP2', 2), ('hex', 'ElementHex1', 1), ('hex', 'ElementHex2', 2),
])
def test_interpolate_reproduces_polynomials(kind, elem, deg):
    """interpolate of the nodal values of a polynomial the element contains
    gives its values and gradients at the mapped quadrature points (hex: a
    trilinear / triquadratic field on an undistorted mesh, where the mapping
    is linear); the DOF values come from the DOF locations."""
    n = {'tri': 6, 'tet': 4, 'hex': 3}[kind]
    x = np.linspace(0, 1, n)
    if kind == 'tri':
        m = O.MeshTri.init_tensor(x, x)
    elif kind == 'tet':
        m = O.MeshTet.init_tensor(x, x, x)
    else:
        m = O.MeshHex.init_tensor(x, x, x)
    if kind != 'hex':
        m = O.jitter(m, 0.2, 1.0 / (n - 1))
    b = O.CellBasis(m, getattr(O, elem)())
    D = m.p.shape[0]
    rng = np.random.default_rng(3)
    c = rng.standard_normal(D + 1)
    C2 = rng.standard_normal((D, D)) if deg == 2 else np.zeros((D, D))
    if kind == 'hex' and deg == 1:
        C2 = np.zeros((D, D))

    def f(X):
        return c[0] + np.einsum('i,i...->...', c[1:], X) + np.einsum('ij,i...,j...->...', C2, X, X)

    def gradf(X):
        return c[1:].reshape((D,) + (1,) * (X.ndim - 1)) + np.einsum('ij,j...->i...', C2 + C2.T, X)

    # DOF locations on the reference cell, checked to be nodal, mapped to
    # every element; w holds f at each DOF (consistent across elements)
    Xloc = np.array(_DOFLOCS[elem], dtype=float).T
    for i in range(b.Nbfun):
        phi_i, _ = b.elem.lbasis(Xloc, i)
        assert np.allclose(phi_i, np.eye(b.Nbfun)[i], atol=1e-14)
    pts = b.mapping.F(Xloc)
    w = np.zeros(b.N)
    for j in range(b.Nbfun):
        w[b.element_dofs[j]] = f(pts[:, :, j])
    val, grad = O.interpolate(b, w)
    xq = b.mapping.F(b.X)
    assert np.abs(val - f(xq)).max() < 1e-12
    assert np.abs(grad - gradf(xq)).max() < 1e-10


def test_interpolate_partition_of_unity():
    x = np.linspace(0, 1, 5)
    m = O.jitter(O.MeshTet.init_tensor(x, x, x), 0.2, 0.25)
    b = O.CellBasis(m, O.ElementTetP2())
    val, grad = O.interpolate(b, np.ones(b.N))
    assert np.abs(val - 1).max() < 1e-14 and np.abs(grad).max() < 1e-12


def test_condense_oracle():
    """condense(D=boundary) and enforce give the same solution, also with
    nonzero boundary values x[D] (the condensed right-hand side carries
    -A_ID x_D); the ex01 maximum is reproduced through condense."""
    from scipy.sparse.linalg import spsolve
    x = np.linspace(0, 1, 33)
    m = O.MeshTri.init_tensor(x, x)
    b = O.CellBasis(m, O.ElementTriP1())
    A = O.bilinear_assemble(O.form_laplace, b)
    f = O.linear_assemble(O.form_unit_load, b)
    D = m.boundary_nodes()
    xb = np.zeros(A.shape[0])
    xb[D] = np.sin(3 * m.p[0, D]) + m.p[1, D]
    AII, bI, xx, I = O.condense(A, f, x=xb, D=D)
    assert AII.shape == (len(I), len(I)) and AII.has_sorted_indices
    u1 = xx.copy()
    u1[I] = spsolve(AII.tocsc(), bI)
    Ae, fe = O.enforce(A, f, x=xb, D=D)
    u2 = spsolve(Ae.tocsc(), fe)
    assert np.abs(u1 - u2).max() < 1e-12
    # I given instead of D: the same system
    AII2, bI2, _, I2 = O.condense(A, f, x=xb, I=I)
    assert np.array_equal(I2, I) and (AII2 != AII).nnz == 0 and np.array_equal(bI2, bI)
    # ex01 through condense
    x = np.linspace(0, 1, 65)
    m = O.MeshTri.init_tensor(x, x)
    b = O.CellBasis(m, O.ElementTriP1())
    A = O.bilinear_assemble(O.form_laplace, b)
    f = O.linear_assemble(O.form_unit_load, b)
    AII, bI, xx, I = O.condense(A, f, I=m.interior_nodes())
    u = xx.copy()
    u[I] = spsolve(AII.tocsc(), bI)
    assert abs(u.max() - 0.073657185490792) < 1e-7


# -- quadrature ---------------------------------------------------------------
@pytest.mark.parametrize('kind,order', [('tri', 2), ('tri', 3), ('tri', 4),
                                        ('tet', 1), ('tet', 2), ('tet', 3),
                                        ('tet', 4), ('hex', 3), ('hex', 5)])
def test_quadrature_exactness(kind, order):
    from math import factorial
    X, W = {'tri': O.quadrature_tri, 'tet': O.quadrature_tet,
            'hex': O.quadrature_hex}[kind](order)
    d = X.shape[0]
    for a in range(order + 1):
        for bb in range(order + 1 - a):
            for c in range(order + 1 - a - bb if d == 3 else 1):
                mono = X[0] ** a * X[1] ** bb * (X[2] ** c if d == 3 else 1.0)
                num = np.dot(W, mono)
                if kind == 'hex':
                    exact = 1.0 / ((a + 1) * (bb + 1) * (c + 1))
                elif d == 2:
                    exact = factorial(a) * factorial(bb) / factorial(a + bb + 2)
                else:
                    exact = (factorial(a) * factorial(bb) * factorial(c)
                             / factorial(a + bb + c + 3))
                # tabulated rules carry 15-16 significant digits
                assert abs(num - exact) < 5e-14, (a, bb, c, num, exact)


# -- reference elements -------------------------------------------------------
NODES = {
    'tri1': np.array([[0., 0.], [1., 0.], [0., 1.]]),
    'tri2': np.array([[0., 0.], [1., 0.], [0., 1.], [.5, 0.], [.5, .5], [0., .5]]),
    'tet1': np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]]),
    'tet2': np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.],
                      [.5, 0., 0.], [.5, .5, 0.], [0., .5, 0.], [0., 0., .5],
                      [.5, 0., .5], [0., .5, .5]]),
    'hex1': np.array([[1., 1., 1.], [1., 1., 0.], [1., 0., 1.], [0., 1., 1.],
                      [1., 0., 0.], [0., 1., 0.], [0., 0., 1.], [0., 0., 0.]]),
    'hex2': O.HEX2_DOFLOCS,
}


@pytest.mark.parametrize('el', list(NODES))
def test_lbasis_nodality_unity_fd(el):
    """tests/test_elements.py:56-329: nodality, partition of unity and
    finite-difference derivatives."""
    e = EL[el]()
    Xn = NODES[el].T
    nb = Xn.shape[1]
    V = np.array([e.lbasis(Xn, i)[0] for i in range(nb)])
    assert np.abs(V - np.eye(nb)).max() < 1e-13
    rng = np.random.default_rng(0)
    d = Xn.shape[0]
    X = rng.uniform(0.05, 0.3, size=(d, 20))
    S = sum(e.lbasis(X, i)[0] for i in range(nb))
    assert np.abs(S - 1).max() < 1e-13
    h = 1e-6
    for i in range(nb):
        _, g = e.lbasis(X, i)
        for k in range(d):
            Xp, Xm = X.copy(), X.copy()
            Xp[k] += h
            Xm[k] -= h
            fd = (e.lbasis(Xp, i)[0] - e.lbasis(Xm, i)[0]) / (2 * h)
            assert np.abs(fd - np.broadcast_to(g[k], fd.shape)).max() < 1e-7


# -- mappings and local matrices ---------------------------------------------
def test_affine_mapping_inverse():
    x = np.linspace(0, 1, 5)
    m = O.jitter(O.MeshTet.init_tensor(x, x, x), 0.2, 0.25)
    mp = O.MappingAffine(m)
    prod = np.einsum('ijk,jlk->ilk', mp.A, mp.invA)
    assert np.abs(prod - np.eye(3)[:, :, None]).max() < 1e-12


def test_isoparametric_constant_jacobian():
    # node 0 (reference (1,1,1)) sits at the lower grid corner and the
    # reference axes x, y, z run along -z, -x, -y of the grid: J = -P
    x = np.linspace(0, 2, 3)
    m = O.MeshHex.init_tensor(x, x, x)
    mp = O.MappingIsoparametric(m)
    X, _ = O.quadrature_hex(5)
    P = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]])
    for i in range(3):
        for j in range(3):
            J = mp.J(i, j, X)
            assert np.abs(J + P[i, j]).max() < 1e-14


def test_p1_local_closed_forms():
    """P1 Laplace K = |detA|/d! G G^T and mass |detA|/d! (1+delta)/((d+1)(d+2))."""
    x = np.linspace(0, 1, 4)
    for m in (O.jitter(O.MeshTri.init_tensor(x, x), 0.2, 1 / 3),
              O.jitter(O.MeshTet.init_tensor(x, x, x), 0.2, 1 / 3)):
        d = m.p.shape[0]
        e = O.ElementTriP1() if d == 2 else O.ElementTetP1()
        b = O.CellBasis(m, e)
        nb = d + 1
        _, K, _ = O.bilinear_coo(O.form_laplace, b)
        _, M, _ = O.bilinear_coo(O.form_mass, b)
        K = K.reshape(nb, nb, -1)
        M = M.reshape(nb, nb, -1)
        mp = b.mapping
        fact = 2.0 if d == 2 else 6.0
        for el in range(m.nelements):
            Bm = np.vstack([-np.ones(d), np.eye(d)])
            G = Bm @ mp.invA[:, :, el]
            Kc = abs(mp.detA[el]) / fact * G @ G.T
            assert np.abs(K[:, :, el] - Kc).max() < 1e-12 * np.abs(Kc).max()
            Mc = abs(mp.detA[el]) / fact * (np.ones((nb, nb)) + np.eye(nb)) / ((d + 1) * (d + 2))
            assert np.abs(M[:, :, el] - Mc).max() < 1e-14


def test_coo_to_csr_semantics():
    """eliminate_zeros before the duplicate sum: a cancellation zero stays,
    an explicit zero goes (coo_data.py:34-36)."""
    A = O.coo_to_csr(np.array([[0, 0, 1, 1, 1, 2], [1, 1, 0, 2, 2, 2]]),
                     np.array([1., -1., 0., 0., 0., 5.]), (3, 3))
    assert A.indptr.tolist() == [0, 1, 1, 2]
    assert A.indices.tolist() == [1, 2]
    assert A.data.tolist() == [0., 5.]
    assert A.indices.dtype == np.int32


# -- the arithmetic the exact-zero pattern depends on -------------------------
def test_numpy_arithmetic_order():
    """einsum accumulates sequentially from 0 with separate multiply/add;
    np.sum over the quadrature axis is 0 + NumPy's pairwise blocking.  The
    device kernels reproduce exactly this."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 3, 300, 11))
    B = rng.standard_normal((3, 11))
    ref = np.einsum('ijkl,il->jkl', A, B)
    seq = np.zeros((3, 300, 11))
    for i in range(3):
        seq = seq + A[i] * B[i][None, None, :]
    assert np.array_equal(ref, seq)

    def pw(a):
        n = len(a)
        if n < 8:
            r = -0.0
            for x in a:
                r = r + x
            return r
        if n <= 128:
            r = list(a[:8])
            i = 8
            while i < n - n % 8:
                for j in range(8):
                    r[j] = r[j] + a[i + j]
                i += 8
            res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
            for k in range(i, n):
                res = res + a[k]
            return res
        n2 = n // 2
        n2 -= n2 % 8
        return pw(a[:n2]) + pw(a[n2:])

    for Q in (3, 4, 6, 11, 27, 64, 343):
        X = rng.standard_normal((200, Q)) * np.exp(5 * rng.standard_normal((200, Q)))
        mine = np.array([0.0 + pw(r) for r in X])
        assert np.array_equal(np.sum(X, axis=1), mine), Q


def test_planted_mistake_is_caught():
    """A plausible oracle mistake (a transposed local matrix, i.e. rows and
    columns swapped) fails the fixture comparison on a non-symmetric case:
    here a perturbed quadrature weight, which changes values in the last
    bits and hence the exact-zero pattern of the unjittered P2 triangle."""
    z = np.load(os.path.join(GOLD, 'small_tri2_n17.npz'))
    m = O.MeshTri(z['p'], z['t'])
    b = O.CellBasis(m, O.ElementTriP2())
    b.dx = b.dx * (1 + 2.0 ** -50)
    A = O.bilinear_assemble(O.form_laplace, b)
    same = (A.nnz == len(z['laplace_indices'])
            and np.array_equal(A.indices, z['laplace_indices']))
    assert not same or not np.array_equal(A.data, z['laplace_data'])


def test_elasticity_s_matrix_identity():
    """The MFMA kernel's formulation: K[(a,c),(b,d)] = lam S_cd[a,b] +
    mu S_dc[a,b] + mu delta_cd tr(S)[a,b], S_kl = G_k^T diag(w|J|) G_l,
    equals the oracle's ddot(C(sym_grad u), sym_grad v) local matrices."""
    x = np.linspace(0, 1, 3)
    m = O.jitter(O.MeshHex.init_tensor(x, x, x), 0.15, 0.5)
    lam, mu = O.lame_parameters(1.0, 0.3)
    vb = O.CellBasis(m, O.ElementVector(O.ElementHex2()), intorder=5)
    _, data, _ = O.bilinear_coo(O.form_elasticity(lam, mu), vb)
    K = data.reshape(81, 81, -1)          # [j trial, i test, e]
    sb = O.CellBasis(m, O.ElementHex2(), intorder=5)
    g = np.array([sb.basis[a][1] for a in range(27)])   # (27, 3, E, Q)
    om = sb.dx                                           # (E, Q)
    S = np.einsum('akeq,bleq,eq->klabe', g, g, om)       # (3,3,27,27,E)
    tr = S[0, 0] + S[1, 1] + S[2, 2]
    Kf = np.zeros_like(K)
    for i in range(81):
        a, c = divmod(i, 3)
        for j in range(81):
            b, d = divmod(j, 3)
            v = lam * S[c, d, a, b] + mu * S[d, c, a, b]
            if c == d:
                v = v + mu * tr[a, b]
            Kf[j, i] = v
    assert np.abs(Kf - K).max() <= 1e-13 * np.abs(K).max()


def test_threaded_assembly_identical():
    """nthreads splits the (j, i) entries over threads (the reference's
    knob); the matrix is the same bit for bit."""
    x = np.linspace(0, 1, 9)
    m = O.jitter(O.MeshTet.init_tensor(x, x, x), 0.2, 1 / 8)
    b = O.CellBasis(m, O.ElementTetP2())
    A1 = O.bilinear_assemble(O.form_laplace, b)
    A2 = O.bilinear_assemble(O.form_laplace, b, nthreads=3)
    assert np.array_equal(A1.indptr, A2.indptr) and np.array_equal(A1.indices, A2.indices)
    assert np.array_equal(A1.data, A2.data)


@pytest.mark.parametrize('name', ['small_vtet1_n5j', 'small_vtri2_n5j'])
def test_vector_laplace_is_expanded_scalar_laplace(name):
    """The reference's vector_laplace on ElementVector(e) equals laplace on e
    expanded to the components (row dim*r + c = scalar row r, columns
    dim*s + c), bit for bit: the relation the device path relies on."""
    z = np.load(os.path.join(GOLD, name + '.npz'))
    meta = json.loads(str(z['meta']))
    M = {'tri': O.MeshTri, 'tet': O.MeshTet}[meta['kind']]
    E = {'tet1': O.ElementTetP1, 'tri2': O.ElementTriP2}[meta['element']]
    m = M(z['p'], z['t'])
    S = O.assemble_named('laplace', O.CellBasis(m, E()))
    dim = m.p.shape[0]
    n = S.shape[0]
    rows, cols, vals = [], [], []
    for c in range(dim):
        coo = S.tocoo()
        rows.append(dim * coo.row + c)
        cols.append(dim * coo.col + c)
        vals.append(coo.data)
    from scipy.sparse import csr_matrix
    V = csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                   shape=(dim * n, dim * n))
    V.sort_indices()
    assert np.array_equal(V.indptr, z['vector_laplace_indptr'])
    assert np.array_equal(V.indices, z['vector_laplace_indices'])
    assert np.array_equal(V.data, z['vector_laplace_data'])
