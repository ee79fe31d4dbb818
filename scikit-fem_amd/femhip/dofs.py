"""Global DOF numbering (``skfem/assembly/dofs.py:264-334``).

Blocks in order: nodal DOFs (``nodal_dofs`` per vertex, interleaved
``nodal_dofs * vertex + k``), then edge DOFs (3-D only), facet DOFs and
interior DOFs, each block numbered in Fortran order.  ``element_dofs`` is
``(nbfun, nelems)`` int32 with rows in the element's local basis order.
"""
import numpy as np


def _block(t2x, per, offset):
    """Rows ``per * t2x[r] + k + offset`` for every local entity r and k."""
    rows = []
    for r in range(t2x.shape[0]):
        for k in range(per):
            rows.append(per * t2x[r] + k + offset)
    return rows


class Dofs:

    def __init__(self, mesh, elem):
        nv = mesh.nvertices
        E = mesh.nelements
        rows = []
        offset = 0
        rows += _block(mesh.t, elem.nodal_dofs, offset)
        offset += elem.nodal_dofs * nv
        self.nodal_dofs = (np.arange(elem.nodal_dofs * nv, dtype=np.int32)
                           .reshape((elem.nodal_dofs, nv), order='F'))
        if elem.dim == 3 and elem.edge_dofs > 0:
            rows += _block(mesh.t2e, elem.edge_dofs, offset)
            offset += elem.edge_dofs * mesh.nedges
        if elem.facet_dofs > 0:
            rows += _block(mesh.t2f, elem.facet_dofs, offset)
            offset += elem.facet_dofs * mesh.nfacets
        if elem.interior_dofs > 0:
            ar = np.arange(E, dtype=np.int64)
            for k in range(elem.interior_dofs):
                rows.append(elem.interior_dofs * ar + k + offset)
            offset += elem.interior_dofs * E
        ed = np.vstack(rows) if rows else np.zeros((0, E))
        if offset >= 2 ** 31 - 1:
            raise OverflowError("number of DOFs exceeds int32")
        self.element_dofs = np.ascontiguousarray(ed.astype(np.int32))
        self.N = int(self.element_dofs.max()) + 1 if self.element_dofs.size else 0
