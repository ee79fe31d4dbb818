"""Quadrature rules on the reference cells.

The triangle and tetrahedron rules are the tabulated rules scikit-fem uses
(``skfem/quadrature.py:80-905``): the same published point/weight values are
kept digit for digit because the exact-zero pattern of the assembled matrix
depends on the last bit of every weight.  Lines and hexahedra use
Gauss-Legendre rules mapped to [0, 1] and their tensor product in Fortran
order (``skfem/quadrature.py:64-74,2839-2844``).
"""
import numpy as np
from numpy.polynomial.legendre import leggauss

_SIXTH = 1.666666666666666666666e-01
_TWO_THIRDS = 6.666666666666666666666e-01

# order -> (points (2, Q), weights (Q,)), published symmetric rules
_TRI_RULES = {
    2: ([[_SIXTH, _TWO_THIRDS, _SIXTH], [_SIXTH, _SIXTH, _TWO_THIRDS]],
        [_SIXTH, _SIXTH, _SIXTH]),
    3: ([[0.333333333333333, 0.2, 0.6, 0.2],
         [0.333333333333333, 0.6, 0.2, 0.2]],
        [-0.28125, 0.260416666666667, 0.260416666666667, 0.260416666666667]),
    4: ([[0.445948490915965, 0.0915762135097699, 0.10810301816807,
          0.816847572980459, 0.445948490915965, 0.091576213509771],
         [0.10810301816807, 0.816847572980459, 0.445948490915965,
          0.091576213509771, 0.445948490915965, 0.0915762135097699]],
        [0.111690794839006, 0.054975871827661, 0.111690794839006,
         0.054975871827661, 0.111690794839006, 0.054975871827661]),
}


def _tet_rules():
    a, c = 0.5854101966249685, 0.1381966011250105
    s = 0.1666666666666667
    p, q = 0.7857142857142857, 0.0714285714285714
    r, t = 0.1005964238332008, 0.3994035761667992
    w4 = [-0.0789333333333333] + [0.0457333333333333] * 4 \
        + [0.1493333333333333] * 6
    return {
        1: (np.array([[.25], [.25], [.25]]), np.array([1 / 6])),
        2: (np.array([[a, c, c, c], [c, c, c, a], [c, c, a, c]]),
            np.array([0.25, 0.25, 0.25, 0.25]) / 6.0),
        3: (np.array([[0.25, 0.5, s, s, s], [0.25, s, s, s, 0.5],
                      [0.25, s, s, 0.5, s]]),
            np.array([-0.8, 0.45, 0.45, 0.45, 0.45]) / 6.0),
        4: (np.array([[0.25, p, q, q, q, r, t, t, t, r, r],
                      [0.25, q, q, q, p, t, r, t, r, t, r],
                      [0.25, q, q, p, q, t, t, r, r, r, t]]),
            np.array(w4) / 6.0),
    }


def quadrature_tri(norder):
    norder = max(int(norder), 2)
    if norder not in _TRI_RULES:
        raise NotImplementedError("The requested order of quadrature"
                                  "is not implemented!")
    X, W = _TRI_RULES[norder]
    return np.array(X, dtype=np.float64), np.array(W, dtype=np.float64)


def quadrature_tet(norder):
    norder = max(int(norder), 1)
    rules = _tet_rules()
    if norder not in rules:
        raise NotImplementedError("The requested order of quadrature"
                                  "is not implemented!")
    return rules[norder]


def quadrature_line(norder):
    """Gauss-Legendre with ceil((n+1)/2) points on [0, 1]."""
    norder = max(int(norder), 2)
    X, W = leggauss(int(np.ceil((norder + 1.0) / 2.0)))
    return np.array([0.5 * X + 0.5]), W / 2.0


def quadrature_hex(norder):
    X, W = quadrature_line(norder)
    A, B, C = np.meshgrid(X, X, X)
    Y = np.vstack((A.flatten(order="F"), B.flatten(order="F"),
                   C.flatten(order="F")))
    A, B, C = np.meshgrid(W, W, W)
    return Y, (A * B * C).flatten(order="F")


def get_quadrature(kind, norder):
    """(X (dim, Q), W (Q,)) for the reference cell ``kind``."""
    if kind == 'tri':
        return quadrature_tri(norder)
    if kind == 'tet':
        return quadrature_tet(norder)
    if kind == 'hex':
        return quadrature_hex(norder)
    if kind == 'line':
        return quadrature_line(norder)
    raise NotImplementedError("The given reference domain type '%s' "
                              "is not supported!" % kind)
