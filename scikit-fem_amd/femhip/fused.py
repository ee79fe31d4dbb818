"""Fused P1 path: plan construction (host C++ via the C ABI) and launch.

The plan depends only on the mesh topology, the node coordinates used for
the Morton order (any order is valid; only locality depends on it) and the
CSR pattern, so one plan serves every later assembly with that pattern --
laplace, kappa_laplace with any kappa that prunes the same entries, and the
adjoint's forward pass.
"""
import ctypes

import numpy as np
import torch

from . import _lib

# LDS budget per workgroup: two workgroups per CU (2 x 80 KB of 160 KB)
DEFAULT_SLOT_CAP = 6144
DEFAULT_NODE_CAP = 1024
BLOCK = 512

_FORM = {'laplace': _lib.LAPLACE, 'mass': _lib.MASS, 'kappa_laplace': _lib.KAPPA_LAPLACE}


def eligible(basis, name):
    e = basis.elem
    return (not e.vector and e.nbs == e.dim + 1 and e.kind in ('tri', 'tet')
            and e.maxdeg == 1 and name in _FORM)


class FusedPlan:
    """Device arrays of one fused plan plus its info struct."""

    def __init__(self, mesh, indptr, indices, slot_cap=DEFAULT_SLOT_CAP,
                 node_cap=DEFAULT_NODE_CAP):
        lib = _lib.load()
        t = np.ascontiguousarray(mesh.t, dtype=np.int32)
        p = np.ascontiguousarray(mesh.p, dtype=np.float64)
        ip = np.ascontiguousarray(indptr, dtype=np.int32)
        ix = np.ascontiguousarray(indices, dtype=np.int32)
        handle = ctypes.c_void_p()
        info = _lib.FusedInfo()
        _lib.check(lib.femhip_fused_plan_build(
            t.shape[0], p.shape[0], t.shape[1], p.shape[1],
            t.ctypes.data_as(ctypes.c_void_p), p.ctypes.data_as(ctypes.c_void_p),
            ip.ctypes.data_as(ctypes.c_void_p), ix.ctypes.data_as(ctypes.c_void_p),
            slot_cap, node_cap, BLOCK, ctypes.byref(handle), ctypes.byref(info)))
        try:
            parts = np.empty(8 * info.nparts, np.int32)
            nodemap = np.empty(info.nlocal, np.int32)
            rows = np.empty(2 * info.nrows, np.int32)
            chunks = np.empty(2 * info.nchunks, np.int32)
            recs = np.empty(info.rec_words * info.nrecs, np.uint32)
            rec_elem = np.empty(info.nrecs, np.int32)
            arrs = (parts, nodemap, rows, chunks, recs, rec_elem)
            _lib.check(lib.femhip_fused_plan_export(
                handle, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))
        finally:
            lib.femhip_fused_plan_free(handle)
        self.info = info
        dev = torch.device('cuda')
        self.parts = torch.from_numpy(parts).to(dev)
        self.nodemap = torch.from_numpy(nodemap).to(dev)
        self.rows = torch.from_numpy(rows).to(dev)
        self.chunks = torch.from_numpy(chunks).to(dev)
        self.recs = torch.from_numpy(recs.view(np.int32)).to(dev)
        self.rec_elem = torch.from_numpy(rec_elem).to(dev)
        self.dev = _lib.FusedDev(self.parts.data_ptr(), self.nodemap.data_ptr(),
                                 self.rows.data_ptr(), self.chunks.data_ptr(),
                                 self.recs.data_ptr(), self.rec_elem.data_ptr(), info)
        self.nnz = int(indptr[-1])

    def stats(self):
        i = self.info
        return dict(parts=i.nparts, records=int(i.nrecs), dup=i.dup,
                    max_colors=i.max_colors, rec_bytes=4 * i.rec_words,
                    max_slots=i.max_slots, max_nodes=i.max_nodes,
                    plan_bytes=int(sum(t.numel() * t.element_size() for t in (
                        self.parts, self.nodemap, self.rows, self.chunks,
                        self.recs, self.rec_elem))))

    def assemble(self, name, p, out, kappa=None):
        """out[nnz] <- CSR values of ``name`` for node coordinates ``p``."""
        lib = _lib.load()
        assert out.numel() == self.nnz and out.dtype == torch.float64
        _lib.check(lib.femhip_fused_assemble(
            ctypes.byref(self.dev), _FORM[name], p.shape[0], _lib.ptr(p), p.shape[1],
            _lib.ptr(kappa), _lib.ptr(out), _lib.stream_ptr()))
        return out
