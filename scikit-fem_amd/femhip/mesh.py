"""Meshes of the accelerated path: first-order triangles, tetrahedra and
hexahedra with scikit-fem's node and element numbering.

``init_tensor`` reproduces ``MeshTri1.init_tensor`` (``mesh_tri_1.py:31-79``,
columns of ``t`` sorted as ``sort_t=True`` does at ``mesh.py:555-556``),
``MeshTet1.init_tensor`` (six Kuhn tetrahedra per cube in six blocks of
``ne``, ``mesh_tet_1.py:327-393``) and ``MeshHex1.init_tensor``
(``mesh_hex_1.py:98-155``) with vectorized index arithmetic.

Edges and facets are numbered as ``Mesh.build_entities`` numbers them
(``mesh.py:1066-1082``): by lexicographic order of the sorted vertex tuple.
Here the tuple is packed into one int64 key (or compared row-wise for
four-vertex facets), which gives the same order as ``np.unique(axis=1)``
at a fraction of the cost on million-element meshes.
"""
import numpy as np

# local entity orderings of the reference cells (skfem/refdom.py)
TRI_FACETS = ((0, 1), (1, 2), (0, 2))
TET_FACETS = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))
TET_EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))
HEX_FACETS = ((0, 1, 4, 2), (0, 2, 6, 3), (0, 3, 5, 1),
              (2, 4, 7, 6), (1, 5, 7, 4), (3, 6, 7, 5))
HEX_EDGES = ((0, 1), (0, 2), (0, 3), (1, 4), (1, 5), (2, 4),
             (2, 6), (3, 5), (3, 6), (4, 7), (5, 7), (6, 7))


def _entities(t, local, nv, keep_order=False):
    """(entities, t2x) for the local vertex tuples ``local`` of every element.

    Entity ids follow the lexicographic order of the sorted vertex tuples;
    ``entities`` holds the sorted tuple, or with ``keep_order`` the tuple in
    the element-local order of its first occurrence (as the reference's
    ``sort=False`` does for hexahedral facets).
    """
    k = len(local[0])
    E = t.shape[1]
    idx = np.concatenate([t[list(ix)] for ix in local], axis=1)  # (k, n*E)
    srt = np.sort(idx, axis=0).astype(np.int64)
    if k <= 3 and float(nv) ** k < 2.0 ** 62:
        key = srt[0]
        for r in range(1, k):
            key = key * nv + srt[r]
        _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    else:
        order = np.lexsort(srt[::-1])
        s = srt[:, order]
        newgrp = np.ones(s.shape[1], dtype=bool)
        newgrp[1:] = np.any(s[:, 1:] != s[:, :-1], axis=0)
        gid = np.cumsum(newgrp) - 1
        inv = np.empty(s.shape[1], dtype=np.int64)
        inv[order] = gid
        first = order[newgrp]
    ent = (idx[:, first] if keep_order else srt[:, first]).astype(np.int32)
    t2x = inv.reshape(len(local), E).astype(np.int64)
    return np.ascontiguousarray(ent), t2x


class Mesh:
    """Common part of the three mesh types: ``p`` (dim, nnodes) float64 and
    ``t`` (nverts, nelems) int32, both C-contiguous."""

    kind = None
    facet_local = None
    edge_local = None
    keep_facet_order = False
    sort_t = False

    def __init__(self, p=None, t=None):
        if p is None:
            p, t = self._default()
        t = np.asarray(t)
        if self.sort_t:
            t = np.sort(t, axis=0)
        self.p = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
        self.t = np.ascontiguousarray(np.asarray(t, dtype=np.int32))
        if self.p.shape[0] != self.dim or self.t.shape[0] != self.nverts:
            raise ValueError("Incompatible p/t shapes for %s" % type(self).__name__)

    # -- sizes -------------------------------------------------------------
    @property
    def doflocs(self):
        return self.p

    @property
    def nelements(self):
        return self.t.shape[1]

    @property
    def nvertices(self):
        return int(self.t.max()) + 1 if self.t.size else 0

    @property
    def nfacets(self):
        return self.facets.shape[1]

    @property
    def nedges(self):
        return self.edges.shape[1]

    def dim_(self):
        return self.dim

    # -- entities ----------------------------------------------------------
    def _init_facets(self):
        self._facets, self._t2f = _entities(self.t, self.facet_local,
                                            max(self.nvertices, 1),
                                            self.keep_facet_order)

    def _init_edges(self):
        if self.edge_local is None:
            raise NotImplementedError
        self._edges, self._t2e = _entities(self.t, self.edge_local,
                                           max(self.nvertices, 1))

    @property
    def facets(self):
        if not hasattr(self, '_facets'):
            self._init_facets()
        return self._facets

    @property
    def t2f(self):
        if not hasattr(self, '_t2f'):
            self._init_facets()
        return self._t2f

    @property
    def edges(self):
        if not hasattr(self, '_edges'):
            self._init_edges()
        return self._edges

    @property
    def t2e(self):
        if not hasattr(self, '_t2e'):
            self._init_edges()
        return self._t2e

    @property
    def f2t(self):
        """(2, nfacets) int32: the elements on either side of every facet
        (``Mesh.build_inverse``, ``mesh.py:1085-1100``).  Reading ``t2f`` in
        C order (local facet, then element), row 0 holds the element of the
        facet's first occurrence and row 1 that of its last; a facet seen
        once (a boundary facet) has -1 in row 1."""
        if not hasattr(self, '_f2t'):
            t2f = np.asarray(self.t2f)
            E = t2f.shape[1]
            flat = t2f.ravel()
            order = np.argsort(flat, kind='stable')   # positions by facet, C order kept
            cnt = np.bincount(flat, minlength=self.nfacets)
            end = np.cumsum(cnt)
            f2t = np.empty((2, self.nfacets), dtype=np.int32)
            f2t[0] = order[end - cnt] % E
            f2t[1] = order[end - 1] % E
            f2t[1, cnt == 1] = -1
            self._f2t = f2t
        return self._f2t

    @property
    def brefdom(self):
        """Reference domain of the facets (``mesh.py:57-59``)."""
        return {'tri': 'line', 'tet': 'tri', 'hex': 'quad'}[self.kind]

    def facets_satisfying(self, test, boundaries_only=False):
        """Facets whose midpoints satisfy ``test`` (``mesh.py:426-456``,
        without the ``normal`` orientation)."""
        midp = self.p[:, self.facets].mean(axis=1)
        facets = np.nonzero(test(midp))[0].astype(np.int32)
        if boundaries_only:
            facets = np.intersect1d(facets, self.boundary_facets())
        return facets

    def normalize_facets(self, facets):
        """Facet indices from a selection (``mesh.py:1291-1329``): None (the
        boundary), an index or index array, a list/tuple/set of selections
        (their sorted union), a callable on facet midpoints, or the name of
        a boundary in ``self.boundaries``."""
        if isinstance(facets, (int, np.integer)):
            return np.array([facets])
        if isinstance(facets, np.ndarray):
            return facets
        if facets is None:
            return self.boundary_facets()
        if isinstance(facets, (tuple, list, set)):
            return np.unique(np.concatenate([self.normalize_facets(f) for f in facets]))
        if callable(facets):
            return self.facets_satisfying(facets)
        if isinstance(facets, str):
            bnd = getattr(self, 'boundaries', None)
            if bnd is not None and facets in bnd:
                return bnd[facets]
            raise ValueError("Boundary '{}' not found.".format(facets))
        raise NotImplementedError

    def boundary_facets(self):
        """Facets that belong to exactly one element (``mesh.py:197-199``)."""
        counts = np.bincount(self.t2f.ravel(), minlength=self.nfacets)
        return np.nonzero(counts == 1)[0].astype(np.int32)

    def boundary_nodes(self):
        """``mesh.py:398-400``."""
        return np.unique(self.facets[:, self.boundary_facets()])

    def interior_nodes(self):
        """``mesh.py:402-405``."""
        mask = np.ones(self.p.shape[1], dtype=bool)
        mask[self.boundary_nodes()] = False
        return np.nonzero(mask)[0]

    def mapping(self):
        """The mesh's default reference mapping (``mesh.py:513-528``):
        ``MappingAffine`` on simplices, ``MappingIsoparametric`` with the
        trilinear element on hexahedra (a descriptor; the kernels evaluate
        it)."""
        from .mapping import MappingAffine, MappingIsoparametric
        if not hasattr(self, '_cached_mapping'):
            if self.kind == 'hex':
                from .element import ElementHex1
                self._cached_mapping = MappingIsoparametric(self, ElementHex1())
            else:
                self._cached_mapping = MappingAffine(self)
        return self._cached_mapping

    def with_p(self, p):
        """Same topology (and cached entities), new node coordinates."""
        out = type(self).__new__(type(self))
        out.__dict__.update(self.__dict__)
        out.p = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
        return out

    def jittered(self, a, h, seed=0):
        """Interior nodes moved by ``U(-a h, a h)`` drawn from
        ``default_rng(seed)`` in the order ``(dim, len(interior_nodes))``."""
        rng = np.random.default_rng(seed)
        I = self.interior_nodes()
        p = self.p.copy()
        p[:, I] += rng.uniform(-a * h, a * h, size=(p.shape[0], len(I)))
        return self.with_p(p)

    def __repr__(self):
        return "%s(%d vertices, %d elements)" % (type(self).__name__,
                                                 self.p.shape[1], self.nelements)


def _tensor_ids(npy, npx, npz=1):
    """Node ids ix[iy, ix, iz] = iy + npy*ix + npy*npx*iz (Fortran order)."""
    return np.arange(npx * npy * npz).reshape(npy, npx, npz, order='F')


class MeshTri(Mesh):
    """First-order triangles (``MeshTri1``); element columns are sorted."""

    kind, dim, nverts = 'tri', 2, 3
    facet_local = TRI_FACETS
    sort_t = True

    @staticmethod
    def _default():
        return (np.array([[0., 0.], [1., 0.], [0., 1.], [1., 1.]]).T,
                np.array([[0, 1, 2], [1, 3, 2]], dtype=np.int32).T)

    @classmethod
    def init_tensor(cls, x, y):
        x, y = np.sort(np.asarray(x, float)), np.sort(np.asarray(y, float))
        npx, npy = len(x), len(y)
        X, Y = np.meshgrid(x, y)
        p = np.vstack((X.flatten('F'), Y.flatten('F')))
        ix = _tensor_ids(npy, npx)[:, :, 0]
        a = ix[:-1, :-1].flatten('F')
        b = ix[1:, :-1].flatten('F')
        c = ix[1:, 1:].flatten('F')
        d = ix[:-1, 1:].flatten('F')
        t = np.vstack((np.concatenate((a, a)), np.concatenate((b, d)),
                       np.concatenate((c, c))))
        return cls(p, t.astype(np.int32))


class MeshTet(Mesh):
    """First-order tetrahedra (``MeshTet1``)."""

    kind, dim, nverts = 'tet', 3, 4
    facet_local = TET_FACETS
    edge_local = TET_EDGES

    @staticmethod
    def _default():
        return (np.array([[0., 0., 0.], [0., 0., 1.], [0., 1., 0.],
                          [1., 0., 0.], [0., 1., 1.], [1., 0., 1.],
                          [1., 1., 0.], [1., 1., 1.]]).T,
                np.array([[0, 1, 2, 3], [3, 5, 1, 7], [2, 3, 6, 7],
                          [2, 3, 1, 7], [1, 2, 4, 7]], dtype=np.int32).T)

    @classmethod
    def init_tensor(cls, x, y, z):
        p, c = _hex_corners(x, y, z)
        kuhn = ((0, 1, 5, 7), (0, 1, 4, 7), (0, 2, 4, 7),
                (0, 3, 5, 7), (0, 2, 6, 7), (0, 3, 6, 7))
        t = np.hstack([c[list(k)] for k in kuhn])
        return cls(p, t.astype(np.int32))


class MeshHex(Mesh):
    """First-order hexahedra (``MeshHex1``); node 0 of every element is the
    reference vertex (1, 1, 1) mapped to the lower grid corner."""

    kind, dim, nverts = 'hex', 3, 8
    facet_local = HEX_FACETS
    edge_local = HEX_EDGES
    keep_facet_order = True

    @staticmethod
    def _default():
        return (np.array([[0., 0., 0.], [0., 0., 1.], [0., 1., 0.],
                          [1., 0., 0.], [0., 1., 1.], [1., 0., 1.],
                          [1., 1., 0.], [1., 1., 1.]]).T,
                np.arange(8, dtype=np.int32)[:, None])

    @classmethod
    def init_tensor(cls, x, y, z):
        p, c = _hex_corners(x, y, z)
        return cls(p, c.astype(np.int32))


def _hex_corners(x, y, z):
    """Nodes and the 8 corner ids of every grid cell in the reference order
    (shared by the tetrahedral and hexahedral constructors)."""
    x, y, z = (np.sort(np.asarray(v, float)) for v in (x, y, z))
    npx, npy, npz = len(x), len(y), len(z)
    X, Y, Z = np.meshgrid(x, y, z)
    p = np.vstack((X.flatten('F'), Y.flatten('F'), Z.flatten('F')))
    ix = _tensor_ids(npy, npx, npz)
    lo, hi = slice(0, -1), slice(1, None)
    sl = ((lo, lo, lo), (hi, lo, lo), (lo, hi, lo), (lo, lo, hi),
          (hi, hi, lo), (hi, lo, hi), (lo, hi, hi), (hi, hi, hi))
    c = np.vstack([ix[s].flatten('F') for s in sl])
    return p, c


MeshTri1, MeshTet1, MeshHex1 = MeshTri, MeshTet, MeshHex
