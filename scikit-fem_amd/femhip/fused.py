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

import os

# LDS budget per workgroup: three workgroups per CU (3 x 53 KB of 160 KB:
# 4096 int64 row accumulators, 768 staged nodes); FEMHIP_FUSED_SLOT_CAP /
# FEMHIP_FUSED_NODE_CAP override it (tuning).
DEFAULT_SLOT_CAP = int(os.environ.get('FEMHIP_FUSED_SLOT_CAP', 4096))
DEFAULT_NODE_CAP = int(os.environ.get('FEMHIP_FUSED_NODE_CAP', 768))
# workgroup size: 512, or 1024 (fixed-point accumulation only)
BLOCK = int(os.environ.get('FEMHIP_FUSED_BLOCK', 512))
# summation mode: 'fixed' (fixed-point LDS integer atomics) or 'rounds'
# (floating-point colour rounds); both deterministic
DEFAULT_ACCUM = os.environ.get('FEMHIP_FUSED_ACCUM', 'fixed')
FX_UNKNOWN = -2 ** 31

_FORM = {'laplace': _lib.LAPLACE, 'mass': _lib.MASS, 'kappa_laplace': _lib.KAPPA_LAPLACE}


def eligible(basis, name):
    e = basis.elem
    return (not e.vector and e.nbs == e.dim + 1 and e.kind in ('tri', 'tet')
            and e.maxdeg == 1 and name in _FORM)


class FusedPlan:
    """Device arrays of one fused plan plus its info struct."""

    def __init__(self, mesh, indptr, indices, slot_cap=DEFAULT_SLOT_CAP,
                 node_cap=DEFAULT_NODE_CAP, accum=None, block=None):
        lib = _lib.load()
        self.accum = accum or DEFAULT_ACCUM
        if self.accum not in ('fixed', 'rounds'):
            raise ValueError("accum must be 'fixed' or 'rounds'")
        block = int(block or BLOCK)
        if block not in (512, 1024) or (block == 1024 and self.accum != 'fixed'):
            raise ValueError("block must be 512, or 1024 with fixed-point accumulation")
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
            slot_cap, node_cap, block, ctypes.byref(handle), ctypes.byref(info)))
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
        self.caps = (slot_cap, node_cap)
        self.info = info
        self.host = dict(parts=parts, nodemap=nodemap, rows=rows, chunks=chunks,
                         recs=recs, rec_elem=rec_elem)
        self.nnz = int(indptr[-1])
        self.nnodes = p.shape[1]
        self.nelems = t.shape[1]
        self.dim = p.shape[0]
        self._upload()

    def _upload(self):
        h = self.host
        dev = torch.device('cuda')
        self.parts = torch.from_numpy(h['parts']).to(dev)
        self.nodemap = torch.from_numpy(h['nodemap']).to(dev)
        self.rows = torch.from_numpy(h['rows']).to(dev)
        self.chunks = torch.from_numpy(h['chunks']).to(dev)
        self.recs = torch.from_numpy(h['recs'].view(np.int32)).to(dev)
        self.rec_elem = torch.from_numpy(h['rec_elem']).to(dev)
        self.scale = None
        if self.accum == 'fixed':
            # exponent pair per (form, partition); owned by the kernel
            self.scale = torch.full((6 * self.info.nparts,), FX_UNKNOWN, dtype=torch.int32,
                                    device=dev)
        self.dev = _lib.FusedDev(self.parts.data_ptr(), self.nodemap.data_ptr(),
                                 self.rows.data_ptr(), self.chunks.data_ptr(),
                                 self.recs.data_ptr(), self.rec_elem.data_ptr(),
                                 None if self.scale is None else self.scale.data_ptr(),
                                 self.info)

    @classmethod
    def concat(cls, plans):
        """One plan over several meshes: node ids, CSR offsets, records and
        element ids are shifted so that the concatenated coordinate array,
        CSR value array and load vector of the meshes can be assembled in a
        single launch."""
        fmt = {(q.info.rec_words, q.info.idbits, q.info.offbits, q.dim) for q in plans}
        if len(fmt) != 1:
            raise NotImplementedError("meshes need the same record format to be batched")
        out = cls.__new__(cls)
        out.accum = plans[0].accum
        parts, nodemap, rows, chunks, recs, rec_elem = [], [], [], [], [], []
        off_node = off_nnz = off_map = off_row = off_chunk = off_rec = off_el = 0
        info = _lib.FusedInfo()
        for f in ('max_slots', 'max_nodes', 'max_rows', 'max_colors'):
            setattr(info, f, max(getattr(q.info, f) for q in plans))
        q0 = plans[0].info
        info.rec_words, info.idbits, info.offbits, info.block = (
            q0.rec_words, q0.idbits, q0.offbits, q0.block)
        self_offsets = []
        for q in plans:
            h = q.host
            pt = h['parts'].reshape(-1, 8).copy()
            pt[:, 0] += off_map
            pt[:, 2] += off_row
            pt[:, 5] += off_chunk
            parts.append(pt.ravel())
            nodemap.append(h['nodemap'] + off_node)
            rw = h['rows'].reshape(-1, 2).copy()
            rw[:, 0] += off_nnz
            rows.append(rw.ravel())
            ch = h['chunks'].reshape(-1, 2).copy()
            ch[:, 0] += off_rec
            chunks.append(ch.ravel())
            recs.append(h['recs'])
            rec_elem.append(h['rec_elem'] + off_el)
            self_offsets.append((off_node, off_nnz, off_el))
            off_node += q.nnodes
            off_nnz += q.nnz
            off_el += q.nelems
            off_map += len(h['nodemap'])
            off_row += len(h['rows']) // 2
            off_chunk += len(h['chunks']) // 2
            off_rec += len(h['rec_elem'])
        out.host = dict(parts=np.concatenate(parts), nodemap=np.concatenate(nodemap),
                        rows=np.concatenate(rows), chunks=np.concatenate(chunks),
                        recs=np.concatenate(recs), rec_elem=np.concatenate(rec_elem))
        info.nparts = sum(q.info.nparts for q in plans)
        info.nchunks = off_chunk
        info.nrows = off_row
        info.nrecs = off_rec
        info.nlocal = off_map
        info.dup = float(off_rec) / max(1, off_el)
        out.info = info
        out.nnz, out.nnodes, out.nelems, out.dim = off_nnz, off_node, off_el, plans[0].dim
        out.offsets = self_offsets
        out._upload()
        return out

    def lds_bytes(self, load=False):
        return int(_lib.load().femhip_fused_lds_bytes(ctypes.byref(self.info), self.dim,
                                                      int(load), int(self.accum == 'fixed')))

    def stats(self):
        i = self.info
        return dict(accum=self.accum, lds_bytes=self.lds_bytes(),
                    parts=i.nparts, records=int(i.nrecs), dup=i.dup,
                    max_colors=i.max_colors, rec_bytes=4 * i.rec_words,
                    max_slots=i.max_slots, max_nodes=i.max_nodes,
                    plan_bytes=int(sum(t.numel() * t.element_size() for t in (
                        self.parts, self.nodemap, self.rows, self.chunks,
                        self.recs, self.rec_elem))))

    def assemble(self, name, p, out, kappa=None, load=None, groups=0):
        """out[nnz] <- CSR values of ``name`` for node coordinates ``p``
        (and load[nnodes] <- the unit-load vector when given)."""
        lib = _lib.load()
        assert out.numel() == self.nnz and out.dtype == torch.float64
        assert p.shape[1] == self.nnodes and p.is_contiguous()
        assert load is None or load.numel() == self.nnodes
        assert kappa is None or kappa.numel() == self.nelems
        _lib.check(lib.femhip_fused_assemble(
            ctypes.byref(self.dev), _FORM[name], p.shape[0], _lib.ptr(p), p.shape[1],
            _lib.ptr(kappa), _lib.ptr(out), _lib.ptr(load), int(groups),
            _lib.stream_ptr()))
        return out


DEFAULT_BLOCK_ROWS = int(os.environ.get('FEMHIP_SWEEP_BLOCK_ROWS', 4096))
DEFAULT_SWEEP_CHUNK = int(os.environ.get('FEMHIP_SWEEP_CHUNK', 512))


def sweep_plan_arrays(mesh, indptr, indices, block_rows=None, chunk=None):
    """Build a layer-sweep plan on the host; returns (info, dict of arrays)."""
    lib = _lib.load()
    t = np.ascontiguousarray(mesh.t, dtype=np.int32)
    p = np.ascontiguousarray(mesh.p, dtype=np.float64)
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    handle = ctypes.c_void_p()
    info = _lib.SweepInfo()
    _lib.check(lib.femhip_sweep_plan_build(
        t.shape[0], p.shape[0], t.shape[1], p.shape[1],
        t.ctypes.data_as(ctypes.c_void_p), p.ctypes.data_as(ctypes.c_void_p),
        ip.ctypes.data_as(ctypes.c_void_p), ix.ctypes.data_as(ctypes.c_void_p),
        int(block_rows or DEFAULT_BLOCK_ROWS), int(chunk or DEFAULT_SWEEP_CHUNK),
        ctypes.byref(handle), ctypes.byref(info)))
    try:
        arrs = dict(blocks=np.empty(4 * info.nblocks, np.int32),
                    levels=np.empty(6 * info.nlevels, np.int32),
                    nodemap=np.empty(info.nlocal, np.int32),
                    rows=np.empty(2 * info.nrows, np.int32),
                    chunks=np.empty(2 * info.nchunks, np.int32),
                    recs=np.empty(info.rec_words * info.nrecs, np.uint32),
                    rec_elem=np.empty(info.nrecs, np.int32))
        _lib.check(lib.femhip_sweep_plan_export(
            handle, *[arrs[k].ctypes.data_as(ctypes.c_void_p) for k in
                      ('blocks', 'levels', 'nodemap', 'rows', 'chunks', 'recs', 'rec_elem')]))
    finally:
        lib.femhip_sweep_plan_free(handle)
    return info, arrs


class SweepPlan:
    """Layer-sweep plan of the fused P1 kernel, in device memory."""

    kind = 'sweep'

    def __init__(self, mesh, indptr, indices, block_rows=None, chunk=None):
        info, h = sweep_plan_arrays(mesh, indptr, indices, block_rows, chunk)
        self.info, self.host = info, h
        self.caps = (int(block_rows or DEFAULT_BLOCK_ROWS), int(chunk or DEFAULT_SWEEP_CHUNK))
        self.nnz = int(indptr[-1])
        self.nnodes = mesh.p.shape[1]
        self.nelems = mesh.t.shape[1]
        self.dim = mesh.p.shape[0]
        dev = torch.device('cuda')
        self.t = {k: torch.from_numpy(v.view(np.int32) if v.dtype == np.uint32 else v).to(dev)
                  for k, v in h.items()}
        self.dev = _lib.SweepDev(*[self.t[k].data_ptr() for k in
                                   ('blocks', 'levels', 'nodemap', 'rows', 'chunks', 'recs',
                                    'rec_elem')], info)

    def stats(self):
        i = self.info
        return dict(variant='sweep', blocks=i.nblocks, levels=i.nlevels, records=int(i.nrecs),
                    dup=i.dup, max_colors=i.max_colors, rec_bytes=4 * i.rec_words,
                    max_level_slots=i.max_level_slots, max_level_nodes=i.max_level_nodes,
                    lds_bytes=int(_lib.load().femhip_sweep_lds_bytes(ctypes.byref(i), self.dim, 0)),
                    nlocal_over_nodes=float(i.nlocal) / self.nnodes,
                    plan_bytes=int(sum(t.numel() * t.element_size() for t in self.t.values())))

    def assemble(self, name, p, out, kappa=None, load=None, groups=0):
        lib = _lib.load()
        assert out.numel() == self.nnz and out.dtype == torch.float64
        assert p.shape[1] == self.nnodes and p.is_contiguous()
        assert load is None or load.numel() == self.nnodes
        assert kappa is None or kappa.numel() == self.nelems
        _lib.check(lib.femhip_sweep_assemble(
            ctypes.byref(self.dev), _FORM[name], p.shape[0], _lib.ptr(p), p.shape[1],
            _lib.ptr(kappa), _lib.ptr(out), _lib.ptr(load), int(groups), _lib.stream_ptr()))
        return out
