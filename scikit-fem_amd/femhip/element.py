"""Reference elements of the accelerated path.

Each element provides the DOF counts that drive the global numbering
(``Element.nodal_dofs`` ... as in ``skfem/element/element.py:44-48``), the
default quadrature degree ``maxdeg`` and ``lbasis(X, i)``: the reference
basis function ``i`` and its reference gradient at points ``X``.  Only the
tables at the quadrature points travel to the device, where the kernels
stage them in LDS / scalar registers.

The polynomials of P1/P2 triangles and tetrahedra and of the trilinear
hexahedron are written with the same operation order as the reference
(``element_tri_p1.py:18-33``, ``element_tri_p2.py:22-48``,
``element_tet_p1.py:19-45``, ``element_tet_p2.py:26-103``,
``element_hex1.py:23-69``) because the exact-zero pattern of the assembled
matrix depends on their last bit.  The triquadratic element is the tensor
product of 1-D quadratic Lagrange polynomials in the reference's node order
(``element_hex2.py:1222-1252``); it spans the same nodal basis, with values
that agree with the reference's generated table to rounding.
"""
import numpy as np

from .hex2_forms import HEX2_FORMS


class Element:
    nodal_dofs = 0
    edge_dofs = 0
    facet_dofs = 0
    interior_dofs = 0
    maxdeg = -1
    kind = None
    dim = None
    nbs = 0          # scalar basis functions per element
    vector = False

    @property
    def nbfun(self):
        return self.nbs * (self.dim if self.vector else 1)

    def __call__(self):
        return self

    def lbasis(self, X, i):
        raise NotImplementedError

    def tables(self, X):
        """(phi (nbs, Q), dphi (nbs, dim, Q)) at the points X (dim, Q)."""
        nq = X.shape[1]
        phi = np.empty((self.nbs, nq))
        dphi = np.empty((self.nbs, self.dim, nq))
        for i in range(self.nbs):
            v, g = self.lbasis(X, i)
            phi[i] = np.broadcast_to(v, (nq,))
            dphi[i] = np.broadcast_to(g, (self.dim, nq))
        return phi, dphi

    def __repr__(self):
        return type(self).__name__ + "()"


class ElementTriP1(Element):
    nodal_dofs, maxdeg, kind, dim, nbs = 1, 1, 'tri', 2, 3
    doflocs = np.array([[0., 0.], [1., 0.], [0., 1.]])

    def lbasis(self, X, i):
        x, y = X
        if i == 0:
            return 1. - x - y, np.array([-1. + 0. * x, -1. + 0. * x])
        if i == 1:
            return x, np.array([1. + 0. * x, 0. * x])
        if i == 2:
            return y, np.array([0. * x, 1. + 0. * x])
        raise ValueError("Index %d out of range" % i)


class ElementTriP2(Element):
    nodal_dofs, facet_dofs, maxdeg, kind, dim, nbs = 1, 1, 2, 'tri', 2, 6
    doflocs = np.array([[0., 0.], [1., 0.], [0., 1.],
                        [.5, 0.], [.5, .5], [0., .5]])

    def lbasis(self, X, i):
        x, y = X
        if i == 0:
            g = -3 + 4. * x + 4. * y
            return (1. - 3. * x - 3. * y + 2. * x ** 2 + 4. * x * y
                    + 2. * y ** 2, np.array([g, g]))
        if i == 1:
            return 2. * x ** 2 - x, np.array([4. * x - 1, 0. * x])
        if i == 2:
            return 2. * y ** 2 - y, np.array([0. * x, 4. * y - 1])
        if i == 3:
            return (4. * x - 4. * x ** 2 - 4. * x * y,
                    np.array([4 - 8. * x - 4. * y, -4. * x]))
        if i == 4:
            return 4. * x * y, np.array([4. * y, 4. * x])
        if i == 5:
            return (4. * y - 4. * x * y - 4. * y ** 2,
                    np.array([-4. * y, 4 - 4. * x - 8. * y]))
        raise ValueError("Index %d out of range" % i)


class ElementTetP1(Element):
    nodal_dofs, maxdeg, kind, dim, nbs = 1, 1, 'tet', 3, 4
    doflocs = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.],
                        [0., 0., 1.]])

    def lbasis(self, X, i):
        x, y, z = X
        one, zero = 1 + 0 * x, 0 * x
        if i == 0:
            m = -1 + 0 * x
            return 1 - x - y - z, np.array([m, m, m])
        if i == 1:
            return x, np.array([one, zero, zero])
        if i == 2:
            return y, np.array([zero, one, zero])
        if i == 3:
            return z, np.array([zero, zero, one])
        raise ValueError("Index %d out of range" % i)


class ElementTetP2(Element):
    nodal_dofs, edge_dofs, maxdeg, kind, dim, nbs = 1, 1, 2, 'tet', 3, 10

    def lbasis(self, X, i):
        x, y, z = X
        if i == 0:
            g = -3. + 4. * x + 4. * y + 4. * z
            return (1. - 3. * x + 2. * x ** 2 - 3. * y + 4. * x * y
                    + 2. * y ** 2 - 3. * z + 4. * x * z + 4. * y * z
                    + 2. * z ** 2, np.array([g, g, g]))
        if i == 1:
            return -1. * x + 2. * x ** 2, np.array([-1 + 4 * x, 0 * x, 0 * x])
        if i == 2:
            return -1. * y + 2. * y ** 2, np.array([0 * x, -1. + 4. * y, 0 * x])
        if i == 3:
            return -1. * z + 2. * z ** 2, np.array([0 * x, 0 * x, -1. + 4. * z])
        if i == 4:
            return (4. * x - 4. * x ** 2 - 4. * x * y - 4 * x * z,
                    np.array([4. - 8. * x - 4. * y - 4. * z, -4. * x, -4. * x]))
        if i == 5:
            return 4. * x * y, np.array([4. * y, 4. * x, 0 * x])
        if i == 6:
            return (0. + 4. * y - 4. * x * y - 4. * y ** 2 - 4. * y * z,
                    np.array([-4. * y, 4. - 4. * x - 8. * y - 4. * z, -4. * y]))
        if i == 7:
            return (0. + 4. * z - 4. * x * z - 4. * y * z - 4. * z ** 2,
                    np.array([-4. * z, -4. * z, 4. - 4. * x - 4. * y - 8. * z]))
        if i == 8:
            return 0. + 4. * x * z, np.array([4. * z, 0 * x, 4 * x])
        if i == 9:
            return 0. + 4. * y * z, np.array([0 * x, 4 * z, 4 * y])
        raise ValueError("Index %d out of range" % i)


class ElementHex1(Element):
    """Trilinear hexahedron; local node 0 is the reference vertex (1,1,1)."""
    nodal_dofs, maxdeg, kind, dim, nbs = 1, 3, 'hex', 3, 8

    def lbasis(self, X, i):
        x, y, z = X
        if i == 0:
            return x * y * z, np.array([y * z, x * z, x * y])
        if i == 1:
            return (x * y * (1 - z),
                    np.array([y * (1 - z), x * (1 - z), -x * y]))
        if i == 2:
            return (x * (1 - y) * z,
                    np.array([(1 - y) * z, -x * z, x * (1 - y)]))
        if i == 3:
            return ((1 - x) * y * z,
                    np.array([-y * z, (1 - x) * z, (1 - x) * y]))
        if i == 4:
            return (x * (1 - y) * (1 - z),
                    np.array([(1 - y) * (1 - z), -x * (1 - z), -x * (1 - y)]))
        if i == 5:
            return ((1 - x) * y * (1 - z),
                    np.array([-y * (1 - z), (1 - x) * (1 - z), -(1 - x) * y]))
        if i == 6:
            return ((1 - x) * (1 - y) * z,
                    np.array([-(1 - y) * z, -(1 - x) * z, (1 - x) * (1 - y)]))
        if i == 7:
            return ((1 - x) * (1 - y) * (1 - z),
                    np.array([-(1 - y) * (1 - z), -(1 - x) * (1 - z),
                              -(1 - x) * (1 - y)]))
        raise ValueError("Index %d out of range" % i)


# node locations of the triquadratic element: vertices, edge midpoints,
# facet centres, cell centre
HEX2_NODES = np.array([
    [1., 1., 1.], [1., 1., 0.], [1., 0., 1.], [0., 1., 1.], [1., 0., 0.],
    [0., 1., 0.], [0., 0., 1.], [0., 0., 0.],
    [1., 1., .5], [1., .5, 1.], [.5, 1., 1.], [1., .5, 0.], [.5, 1., 0.],
    [1., 0., .5], [.5, 0., 1.], [0., 1., .5], [0., .5, 1.], [.5, 0., 0.],
    [0., .5, 0.], [0., 0., .5],
    [1., .5, .5], [.5, .5, 1.], [.5, 1., .5], [.5, 0., .5], [.5, .5, 0.],
    [0., .5, .5],
    [.5, .5, .5]])


class ElementHex2(Element):
    """Triquadratic hexahedron (27 nodes), evaluated in the reference's
    operation order (``element_hex2.py:11-1260``; see ``hex2_forms``)."""
    nodal_dofs = edge_dofs = facet_dofs = interior_dofs = 1
    maxdeg, kind, dim, nbs = 6, 'hex', 3, 27
    doflocs = HEX2_NODES

    def lbasis(self, X, i):
        if not 0 <= i < 27:
            raise ValueError("Index %d out of range" % i)
        x, y, z = X
        v, g = HEX2_FORMS[i](x, y, z)
        return v, np.array(g)


class ElementVector(Element):
    """The same scalar element in every coordinate direction; local function
    ``i`` is component ``i % dim`` of scalar function ``i // dim``
    (``element_vector.py:36-48``)."""

    vector = True

    def __init__(self, elem, dim=None):
        self.elem = elem
        self.dim = elem.dim if dim is None else dim
        if self.dim != elem.dim:
            raise NotImplementedError("ElementVector with dim != elem.dim")
        self.kind = elem.kind
        self.maxdeg = elem.maxdeg
        self.nbs = elem.nbs
        self.nodal_dofs = elem.nodal_dofs * self.dim
        self.facet_dofs = elem.facet_dofs * self.dim
        self.interior_dofs = elem.interior_dofs * self.dim
        self.edge_dofs = elem.edge_dofs * self.dim

    def lbasis(self, X, i):
        return self.elem.lbasis(X, i)

    def __repr__(self):
        return "ElementVector(%r)" % (self.elem,)
