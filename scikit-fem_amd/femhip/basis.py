"""``CellBasis``: mesh + element + quadrature, the input of every assembly.

Unlike the reference (``cell_basis.py:42-106``), which precomputes the
global basis values and gradients at every (element, quadrature point) on
the host, this basis is lazy: it keeps the mesh arrays, the DOF numbering
and the reference tables at the quadrature points.  The device copies
(int32 connectivity, SoA float64 coordinates, tables) are made once per
basis on first use and reused by every assembly on it.
"""
import ctypes

import numpy as np

from .dofs import Dofs
from .mapping import check_mapping
from .quadrature import get_quadrature
from .element import ElementHex1


class DiscreteField:
    """A field at the quadrature points (``skfem.element.DiscreteField``):
    ``value`` (E, Q) and ``grad`` (dim, E, Q); NumPy arrays from
    ``CellBasis.interpolate``, torch tensors on the device from
    ``CellBasis.interpolate_device``."""

    def __init__(self, value, grad=None):
        self.value, self.grad = value, grad

    @property
    def shape(self):
        return self.value.shape

    def astuple(self):
        return (self.value, self.grad)

    def __array__(self, dtype=None, copy=None):
        v = self.value
        if not isinstance(v, np.ndarray):
            v = v.cpu().numpy()
        return np.asarray(v, dtype=dtype)

    def __getitem__(self, key):
        """Indexing reads the value, as the reference's ndarray subclass
        does (``discrete_field.py:55-56``): ``w.x[0]`` is the x component."""
        return self.value[key]


def normalize_elements(mesh, elements):
    """Element indices from a criterion (``Mesh.normalize_elements``,
    ``mesh.py:1331-1360``): True (all), an index, an index array (taken as
    given), a boolean mask over the elements, a function of the element
    centroids returning such a mask, or a list/tuple/set of these (their
    sorted union)."""
    if isinstance(elements, bool) and elements:
        return np.arange(mesh.nelements, dtype=np.int32)
    if isinstance(elements, (int, np.integer)) and not isinstance(elements, bool):
        return np.array([elements])
    if isinstance(elements, np.ndarray):
        if elements.dtype == bool:
            if elements.shape != (mesh.nelements,):
                raise ValueError("element mask of wrong size")
            return np.flatnonzero(elements)
        return elements
    if callable(elements):
        c = np.asarray(mesh.p)[:, mesh.t].mean(axis=1)
        return np.flatnonzero(np.asarray(elements(c), dtype=bool))
    if isinstance(elements, (tuple, list, set)):
        return np.unique(np.concatenate([normalize_elements(mesh, e) for e in elements]))
    raise NotImplementedError("element criterion %r" % (type(elements),))


class ElementSubset:
    """The elements a ``CellBasis(elements=...)`` covers, as the device
    layer sees a mesh: the same nodes and coordinates, the subset's
    connectivity (other attributes are the full mesh's)."""

    def __init__(self, mesh, tind):
        self.parent = mesh
        self.tind = np.asarray(tind)
        self.kind, self.dim, self.nverts = mesh.kind, mesh.dim, mesh.nverts
        self.p = mesh.p
        self.t = np.ascontiguousarray(np.asarray(mesh.t)[:, self.tind])

    @property
    def nelements(self):
        return self.t.shape[1]

    def __getattr__(self, name):
        return getattr(self.parent, name)


class CellBasis:

    def __init__(self, mesh, elem, mapping=None, intorder=None,
                 elements=None, quadrature=None, dofs=None,
                 disable_doflocs=False):
        if mesh.kind != elem.kind:
            raise ValueError("Incompatible Mesh and Element.")
        check_mapping(mapping, mesh)
        self.mesh = mesh
        self.elem = elem
        self.mapping = mesh.mapping() if mapping is None else mapping
        self.dofs = Dofs(mesh, elem) if dofs is None else dofs
        if quadrature is not None:
            self.X, self.W = (np.asarray(quadrature[0], dtype=np.float64),
                              np.asarray(quadrature[1], dtype=np.float64))
        else:
            order = intorder if intorder is not None else 2 * elem.maxdeg
            self.X, self.W = get_quadrature(mesh.kind, order)
        # elements= (cell_basis.py:94-99): the DOF numbering stays the full
        # mesh's, the assembly runs over the selected elements only
        if elements is None:
            self.tind = None
            self.cell_mesh = mesh
        else:
            self.tind = normalize_elements(mesh, elements)
            self.cell_mesh = ElementSubset(mesh, self.tind)
        self._ed_cache = (None, None)
        self.nelems = self.cell_mesh.nelements
        self._device = None
        self._scalar = None

    @property
    def N(self):
        return self.dofs.N

    @property
    def Nbfun(self):
        return self.element_dofs.shape[0]

    @property
    def element_dofs(self):
        ed = self.dofs.element_dofs
        if self.tind is None:
            return ed
        if self._ed_cache[0] is not ed:
            self._ed_cache = (ed, np.ascontiguousarray(ed[:, self.tind]))
        return self._ed_cache[1]

    @property
    def nodal_dofs(self):
        return self.dofs.nodal_dofs

    def tables(self):
        """Reference tables at the quadrature points:
        (phi (nbs, Q), dphi (nbs, dim, Q), gdphi (8, 3, Q) or None)."""
        phi, dphi = self.elem.tables(self.X)
        gdphi = None
        if self.mesh.kind == 'hex':
            _, gdphi = ElementHex1().tables(self.X)
        return phi, dphi, gdphi

    def zeros(self):
        return np.zeros(self.N)

    def interpolate_device(self, w, grad=True):
        """``interpolate`` with the result left on the device: a
        ``DiscreteField`` of torch tensors (value (E, Q), grad (dim, E, Q))."""
        import torch
        from . import _lib
        if w.shape[0] != self.N:
            raise ValueError("Input array has wrong size.")
        if self.elem.vector:
            return self._interpolate_vector(w, grad)
        d = self.device()
        wt = torch.as_tensor(w, dtype=torch.float64, device=d.p.device).contiguous()
        E, Q, D = d.E, d.Q, d.dim
        value = torch.empty((E, Q), dtype=torch.float64, device=d.p.device)
        gr = torch.empty((D, E, Q), dtype=torch.float64, device=d.p.device) if grad else None
        G = d.geometry() if grad else None
        _lib.check(d.lib.femhip_interpolate(
            ctypes.byref(d.geom), ctypes.byref(d.space), _lib.ptr(G), _lib.ptr(wt),
            _lib.ptr(value), _lib.ptr(gr), _lib.stream_ptr()))
        return DiscreteField(value, gr)

    def _interpolate_vector(self, w, grad):
        """ElementVector: the reference sums the interleaved functions in
        local order, each adding to its own component only
        (abstract_basis.py:289-316, element_vector.py:36-48), i.e. every
        component is the scalar interpolation of the DOFs c, c + dim, ...;
        value (dim, E, Q), grad (dim, dim, E, Q)."""
        import torch
        if self._scalar is None:
            self._scalar = CellBasis(self.mesh, self.elem.elem, quadrature=(self.X, self.W),
                                     elements=self.tind)
        dim = self.elem.dim
        w = np.asarray(w, dtype=np.float64) if not isinstance(w, torch.Tensor) else w
        fs = [self._scalar.interpolate_device(w[c::dim], grad=grad) for c in range(dim)]
        value = torch.stack([f.value for f in fs])
        gr = torch.stack([f.grad for f in fs]) if grad else None
        return DiscreteField(value, gr)

    def interpolate(self, w):
        """A DOF vector at the quadrature points (``abstract_basis.py:271-322``):
        ``DiscreteField`` with ``value`` (E, Q) and ``grad`` (dim, E, Q), equal
        to the reference's bit for bit (computed on the device)."""
        f = self.interpolate_device(np.asarray(w, dtype=np.float64))
        return DiscreteField(f.value.cpu().numpy(), f.grad.cpu().numpy())

    def with_element(self, elem):
        return type(self)(self.mesh, elem, quadrature=(self.X, self.W), elements=self.tind)

    def global_coordinates(self):
        """Coordinates of the quadrature points, ``DiscreteField`` with
        ``value`` (dim, E, Q) (``cell_basis.py:129-134``; bit-identical to the
        reference's mapping ``F``), computed on the device."""
        return DiscreteField(self.device().global_coordinates().cpu().numpy())

    def mesh_parameters(self):
        """``|detDF| ** (1 / dim)`` at the quadrature points, (E, Q)
        (``cell_basis.py:136-141``), computed on the device."""
        return DiscreteField(self.device().mesh_parameters().cpu().numpy())

    def default_parameters(self):
        """``{'x': global_coordinates(), 'h': mesh_parameters()}``
        (``cell_basis.py:124-127``)."""
        return {'x': self.global_coordinates(), 'h': self.mesh_parameters()}

    def device(self):
        """Device copy of this basis (created on first use)."""
        if self._device is None:
            from .device import DeviceBasis
            self._device = DeviceBasis(self)
        return self._device

    def __repr__(self):
        return "CellBasis(%r, %r, %d quadrature points)" % (
            self.mesh, self.elem, self.W.shape[0])


Basis = CellBasis


class FacetBasis:
    """Basis on a set of mesh facets, by default the boundary
    (``facet_basis.py:23-110``), for boundary forms on triangles and
    tetrahedra with P1/P2 elements.

    As ``CellBasis`` it keeps the index data only: ``find`` (the facets),
    ``tind`` (the element on the chosen ``side``, ``f2t[side, find]``),
    ``element_dofs`` (the owning elements' DOFs, ``abstract_basis.py:110-117``)
    and the facet quadrature on ``mesh.brefdom``.  The facet points, their
    pull-back into the owning elements and the local entries are computed on
    the device (``femhip_facet_local``).  A facet without an element on
    ``side`` (side 1 of a boundary facet) raises ``ValueError``.
    """

    def __init__(self, mesh, elem, mapping=None, intorder=None, quadrature=None,
                 facets=None, dofs=None, side=0, disable_doflocs=False):
        if mesh.kind != elem.kind:
            raise ValueError("Incompatible Mesh and Element.")
        check_mapping(mapping, mesh)
        if mesh.kind not in ('tri', 'tet') or elem.vector or elem.maxdeg not in (1, 2):
            raise NotImplementedError("facet bases: P1/P2 triangles and tetrahedra")
        if side not in (0, 1):
            raise ValueError("side is 0 or 1")
        self.mesh = mesh
        self.elem = elem
        self.dofs = Dofs(mesh, elem) if dofs is None else dofs
        if quadrature is not None:
            self.X, self.W = (np.asarray(quadrature[0], dtype=np.float64),
                              np.asarray(quadrature[1], dtype=np.float64))
        else:
            order = intorder if intorder is not None else 2 * elem.maxdeg
            self.X, self.W = get_quadrature(mesh.brefdom, order)
        find = mesh.normalize_facets(facets)
        if find.dtype == bool:
            find = np.nonzero(find)[0]
        self.find = np.ascontiguousarray(find, dtype=np.int32)
        self.side = side
        f2t = mesh.f2t
        self.tind = f2t[side, self.find]
        self.tind_normals = f2t[0, self.find]
        if np.any(self.tind < 0):
            raise ValueError("facets without an element on side %d" % side)
        self.nelems = len(self.find)
        self._element_dofs = None
        self._device = None

    @property
    def N(self):
        return self.dofs.N

    @property
    def Nbfun(self):
        return self.dofs.element_dofs.shape[0]

    @property
    def element_dofs(self):
        if self._element_dofs is None:
            self._element_dofs = self.dofs.element_dofs[:, self.tind]
        return self._element_dofs

    @property
    def nodal_dofs(self):
        return self.dofs.nodal_dofs

    def facet_local_index(self):
        """Local index (``RefTri``/``RefTet`` facet order) of every facet in
        its ``tind_normals`` element."""
        t2f = self.mesh.t2f[:, self.tind_normals]
        return np.argmax(t2f == self.find[None, :], axis=0).astype(np.int32)

    def interpolate_device(self, w, grad=True):
        """``interpolate`` with the result left on the device (a
        ``DiscreteField`` of torch tensors)."""
        import torch
        if w.shape[0] != self.N:
            raise ValueError("Input array has wrong size.")
        d = self.device()
        wt = torch.as_tensor(w, dtype=torch.float64, device=d.p.device).contiguous()
        value, gr = d.interpolate(wt, grad)
        return DiscreteField(value, gr)

    def interpolate(self, w):
        """A DOF vector at the facet quadrature points
        (``abstract_basis.py:271-322`` on a FacetBasis): ``DiscreteField``
        with ``value`` (nfacets, Q) and ``grad`` (dim, nfacets, Q), in the
        reference's operation order, computed on the device."""
        f = self.interpolate_device(np.asarray(w, dtype=np.float64))
        return DiscreteField(f.value.cpu().numpy(), f.grad.cpu().numpy())

    def global_coordinates(self):
        """Facet quadrature points in global coordinates, (dim, nfacets, Q)
        (``facet_basis.py:126-131``)."""
        return DiscreteField(self.device().global_coordinates()[0].cpu().numpy())

    def mesh_parameters(self):
        """``|detB| ** (1 / (dim - 1))`` at the facet quadrature points
        (``facet_basis.py:133-139``)."""
        return DiscreteField(self.device().global_coordinates()[1].cpu().numpy())

    def default_parameters(self):
        """``{'x', 'h', 'n'}`` (``facet_basis.py:118-124``)."""
        return {'x': self.global_coordinates(), 'h': self.mesh_parameters(),
                'n': self.normals}

    @property
    def normals(self):
        """Outward unit normals seen from ``f2t[0]``
        (``mapping_affine.py:255-281``), a ``DiscreteField`` with ``value``
        (dim, nfacets, Q), computed on the device."""
        n = self.device().normals()
        return DiscreteField(n.cpu().numpy())

    def zeros(self):
        return np.zeros(self.N)

    def device(self):
        if self._device is None:
            from .device import FacetDeviceBasis
            self._device = FacetDeviceBasis(self)
        return self._device

    def __repr__(self):
        return "FacetBasis(%r, %r, %d facets, %d quadrature points)" % (
            self.mesh, self.elem, self.nelems, self.W.shape[0])
