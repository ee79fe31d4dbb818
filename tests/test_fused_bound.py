"""The rounding bound behind the fused kernel's changed-input check.

The fused kernel queues an entry as a suspect when |K_ij| <= tau with
    tau = 2^-43 a^(2d-2) / D (1 + 10 a^d / D)
(a = the largest |component| of the edge vectors A = p_k - p_0, D = |det A|),
and rechecks only suspects in the reference's arithmetic.  That is sound
when each of the two arithmetics is within tau / 2 of the exact entry of
the same (rounded) edge vectors.  Here the reference's arithmetic (the
oracle's NumPy restatement of MappingAffine + einsum + the quadrature sum)
is compared with the exact rational value of every entry on elements chosen
to stress the bound: unjittered Kuhn tetrahedra and right triangles (exact
zeros), jittered ones, strongly anisotropic ones, slivers and coordinates
far from unit scale.  The measured errors stay below 1/1000 of tau / 2 on
well-shaped elements (the factor 2^10 above 2^-53 is the margin for the
fused arithmetic's own rounding).  A planted mistake -- the bound without
its 1 / D factor, i.e. in the wrong units -- must fail on the mesh scaled
by 1e-5."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import skfem_np as O


def _meshes(kind):
    rng = np.random.default_rng(7)
    x = np.linspace(0, 1, 4)
    base = O.MeshTri.init_tensor(x, x) if kind == 'tri' else O.MeshTet.init_tensor(x, x, x)
    out = [('unjittered', base), ('jittered', O.jitter(base, 0.2, 1 / 3))]
    aniso = np.array([1e3, 1e-2, 7.0])[:base.p.shape[0], None]
    out.append(('anisotropic', type(base)(base.p * aniso, base.t)))
    for sc in (1e-5, 1e5):
        out.append(('scale %g' % sc, type(base)(O.jitter(base, 0.2, 1 / 3).p * sc, base.t)))
    # slivers: one vertex of every element pulled towards its opposite facet
    t = base.t
    p = [base.p[:, t[k]] for k in range(t.shape[0])]
    dim = base.p.shape[0]
    P = np.hstack(p)                  # every element on its own vertices
    E = t.shape[1]
    T = np.vstack([np.arange(E) + k * E for k in range(t.shape[0])])
    opp = sum(p[k] for k in range(1, t.shape[0])) / (t.shape[0] - 1)
    for frac in (1e-3, 1e-6):
        Q = P.copy()
        Q[:, :E] = opp + frac * (p[0] - opp) + rng.uniform(-1e-9, 1e-9, (dim, E))
        out.append(('sliver %g' % frac, type(base)(Q, T)))
    return out


def _exact_entries(p, t, W):
    """Exact K_ab (rational) of the P1 Laplace matrix of every element for
    the float64 edge vectors A[:, j] = p[t[j+1]] - p[t[0]]."""
    dim = p.shape[0]
    wsum = sum(Fraction(float(w)) for w in W)
    out = []
    for e in range(t.shape[1]):
        A = [[Fraction(float(p[i, t[j + 1, e]] - p[i, t[0, e]])) for j in range(dim)]
             for i in range(dim)]
        if dim == 2:
            det = A[0][0] * A[1][1] - A[0][1] * A[1][0]
            inv = [[A[1][1] / det, -A[0][1] / det], [-A[1][0] / det, A[0][0] / det]]
        else:
            def cof(r, c):
                rr = [i for i in range(3) if i != r]
                cc = [j for j in range(3) if j != c]
                m = A[rr[0]][cc[0]] * A[rr[1]][cc[1]] - A[rr[0]][cc[1]] * A[rr[1]][cc[0]]
                return m if (r + c) % 2 == 0 else -m
            det = sum(A[0][j] * cof(0, j) for j in range(3))
            inv = [[cof(j, i) / det for j in range(3)] for i in range(3)]
        # grad phi_k = invA^T dphi_k: rows of invA, phi_0 = -(sum)
        g = [[-sum(inv[i][j] for i in range(dim)) for j in range(dim)]]
        g += [[inv[k][j] for j in range(dim)] for k in range(dim)]
        K = [[abs(det) * wsum * sum(g[a][d] * g[b][d] for d in range(dim))
              for b in range(dim + 1)] for a in range(dim + 1)]
        out.append((K, A, abs(det)))
    return out


def _tau(A, D, dim, planted=False):
    a = max(abs(float(x)) for row in A for x in row)
    D = float(D)
    inv = 1.0 if planted else 1.0 / D       # planted: the 1 / D factor lost
    if dim == 3:
        return 2.0 ** -43 * a ** 4 * inv * (1 + 10 * a ** 3 / D)
    return 2.0 ** -43 * a ** 2 * inv * (1 + 10 * a ** 2 / D)


@pytest.mark.parametrize('kind', ['tri', 'tet'])
def test_reference_arithmetic_within_half_the_bound(kind):
    worst = 0.0
    nzero = 0
    for name, m in _meshes(kind):
        e = O.ElementTriP1() if kind == 'tri' else O.ElementTetP1()
        b = O.CellBasis(m, e)
        idx, data, shape = O.bilinear_coo(O.form_laplace, b)
        nb = m.t.shape[0]
        K = data.reshape(nb, nb, -1)           # [j, i, e]
        ex = _exact_entries(m.p, m.t, b.W)
        for el, (Kx, A, D) in enumerate(ex):
            tau = _tau(A, D, m.p.shape[0])
            for i in range(nb):
                for j in range(nb):
                    ref = float(K[j, i, el])
                    err = abs(Fraction(ref) - Kx[i][j])
                    nzero += ref == 0.0
                    assert err <= Fraction(tau) / 2, (name, el, i, j, float(err), tau)
                    worst = max(worst, float(err) / tau)
    assert nzero > 0, "the cases include exact zeros"
    assert worst > 0


@pytest.mark.parametrize('kind', ['tri', 'tet'])
def test_planted_bound_in_wrong_units_fails(kind):
    """The same check with the 1 / D factor dropped: the mesh scaled by
    1e-5 violates it, so the check above has teeth."""
    failed = False
    for name, m in _meshes(kind):
        if name != 'scale 1e-05':
            continue
        e = O.ElementTriP1() if kind == 'tri' else O.ElementTetP1()
        b = O.CellBasis(m, e)
        idx, data, shape = O.bilinear_coo(O.form_laplace, b)
        nb = m.t.shape[0]
        K = data.reshape(nb, nb, -1)
        for el, (Kx, A, D) in enumerate(_exact_entries(m.p, m.t, b.W)):
            tau = _tau(A, D, m.p.shape[0], planted=True)
            for i in range(nb):
                for j in range(nb):
                    if abs(Fraction(float(K[j, i, el])) - Kx[i][j]) > Fraction(tau) / 2:
                        failed = True
    assert failed
