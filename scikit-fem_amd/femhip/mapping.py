"""Reference mappings as descriptors (``skfem.mapping``).

The device kernels evaluate the mesh's default mapping themselves
(``MappingAffine`` on triangles and tetrahedra, ``mapping_affine.py:6-232``;
``MappingIsoparametric`` with the mesh's own trilinear element on
hexahedra, ``mapping_isoparametric.py:12-226``), so these classes only name
a mapping: ``Mesh.mapping()`` returns the default one
(``mesh.py:513-528``) and ``CellBasis(mapping=...)`` accepts a mapping that
is the mesh's default.  Other mappings (a quadratic geometry, say) are not
accelerated and raise ``NotImplementedError``.
"""


class Mapping:
    """A mapping of the elements of ``mesh``."""

    def __init__(self, mesh):
        self.mesh = mesh

    def same_as_default(self, mesh):
        """True when this is ``mesh``'s default mapping."""
        raise NotImplementedError


class MappingAffine(Mapping):
    """``x = A x_ref + b`` per simplex (``mapping_affine.py:6-232``)."""

    def same_as_default(self, mesh):
        return mesh.kind in ('tri', 'tet') and _same_mesh(self.mesh, mesh)


class MappingIsoparametric(Mapping):
    """``x = sum_i x_i phi_i(x_ref)`` over the nodes of ``elem``
    (``mapping_isoparametric.py:12-226``); ``bndelem`` is accepted for the
    reference's signature."""

    def __init__(self, mesh, elem, bndelem=None):
        super().__init__(mesh)
        self.elem, self.bndelem = elem, bndelem

    def same_as_default(self, mesh):
        from .element import ElementHex1
        return (mesh.kind == 'hex' and isinstance(self.elem, ElementHex1)
                and _same_mesh(self.mesh, mesh))


def _same_mesh(a, b):
    import numpy as np
    return a is b or (a.kind == b.kind and np.array_equal(a.p, b.p)
                      and np.array_equal(a.t, b.t))


def check_mapping(mapping, mesh):
    """``mapping`` for a basis on ``mesh``: None or the mesh's default one,
    else ``NotImplementedError``."""
    if mapping is None:
        return
    if not isinstance(mapping, Mapping) or not mapping.same_as_default(mesh):
        raise NotImplementedError("only the mesh's default mapping is accelerated "
                                  "(MappingAffine on simplices, MappingIsoparametric "
                                  "with ElementHex1 on hexahedra)")
