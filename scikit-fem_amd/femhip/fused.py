"""Fused P1 path: plan construction (host C++ via the C ABI) and launch.

The plan depends only on the mesh topology, the node coordinates used for
the spatial partition (any coordinates give a valid plan; only locality
depends on them) and the CSR pattern, so one plan serves every later
assembly with that pattern -- laplace, kappa_laplace with any kappa that
prunes the same entries, the adjoint's forward pass.  Whether a call's
inputs prune the same entries is checked during the call itself
(``check``): entries within the rounding bound of zero are recomputed in
the reference's arithmetic and compared with the plan's zero set.
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib

# LDS budget per workgroup: three workgroups per CU (3 x 53 KB of 160 KB:
# 4096 int64 row accumulators, 768 staged nodes); FEMHIP_FUSED_SLOT_CAP /
# FEMHIP_FUSED_NODE_CAP override it (tuning).
DEFAULT_SLOT_CAP = int(os.environ.get('FEMHIP_FUSED_SLOT_CAP', 4096))
DEFAULT_NODE_CAP = int(os.environ.get('FEMHIP_FUSED_NODE_CAP', 768))
# workgroup size (elements per chunk): 512 or 1024
BLOCK = int(os.environ.get('FEMHIP_FUSED_BLOCK', 512))
# record order inside a partition: 0 = colour classes, 1 = element order
DEFAULT_ORDER = int(os.environ.get('FEMHIP_FUSED_ORDER', 1))

_FORM = {'laplace': _lib.LAPLACE, 'mass': _lib.MASS, 'kappa_laplace': _lib.KAPPA_LAPLACE}
STATE_BYTES = ctypes.sizeof(_lib.FusedState)


def eligible(basis, name):
    e = basis.elem
    return (not e.vector and e.nbs == e.dim + 1 and e.kind in ('tri', 'tet')
            and e.maxdeg == 1 and name in _FORM)


class FusedPlan:
    """Device arrays of one fused plan, its info struct, the per-form
    accumulation states and the suspect queue of the validation."""

    def __init__(self, mesh, indptr, indices, slot_cap=DEFAULT_SLOT_CAP,
                 node_cap=DEFAULT_NODE_CAP, block=None, order=None):
        lib = _lib.load()
        block = int(block or BLOCK)
        if block not in (512, 1024):
            raise ValueError("block must be 512 or 1024")
        order = DEFAULT_ORDER if order is None else int(order)
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
            int(slot_cap), int(node_cap), block, order, ctypes.byref(handle),
            ctypes.byref(info)))
        try:
            parts = np.empty(8 * info.nparts, np.int32)
            nodemap = np.empty(info.nlocal, np.int32)
            rows = np.empty(2 * info.nrows, np.int32)
            recs = np.empty(info.rec_words * info.nrecs, np.uint32)
            rec_elem = np.empty(info.nrecs, np.int32)
            arrs = (parts, nodemap, rows, recs, rec_elem)
            _lib.check(lib.femhip_fused_plan_export(
                handle, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))
        finally:
            lib.femhip_fused_plan_free(handle)
        self.caps = (slot_cap, node_cap)
        self.order = order
        self.groups = 0          # launch grid (the autotuner may set it)
        self.info = info
        self.host = dict(parts=parts, nodemap=nodemap, rows=rows, recs=recs,
                         rec_elem=rec_elem)
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
        self.recs = torch.from_numpy(h['recs'].view(np.int32)).to(dev)
        self.rec_elem = torch.from_numpy(h['rec_elem']).to(dev)
        self.states = {}
        self.queue = None
        self.verdict = torch.zeros(4, dtype=torch.int32, device=dev)
        self.dev = _lib.FusedDev(self.parts.data_ptr(), self.nodemap.data_ptr(),
                                 self.rows.data_ptr(), self.recs.data_ptr(),
                                 self.rec_elem.data_ptr(), self.info)

    def state(self, name):
        """Accumulation state of form ``name`` (device, zeroed at creation)."""
        st = self.states.get(name)
        if st is None:
            st = torch.zeros(STATE_BYTES, dtype=torch.uint8, device=self.parts.device)
            self.states[name] = st
        return st

    @classmethod
    def concat(cls, plans):
        """One plan over several meshes: node ids, CSR offsets, records and
        element ids are shifted so that the concatenated coordinate array,
        CSR value array and load vector of the meshes can be assembled in a
        single launch."""
        fmt = {(q.info.rec_words, q.info.idbits, q.info.offbits, q.dim, q.info.block)
               for q in plans}
        if len(fmt) != 1:
            raise NotImplementedError("meshes need the same record format to be batched")
        out = cls.__new__(cls)
        out.order = plans[0].order
        out.groups = 0
        parts, nodemap, rows, recs, rec_elem = [], [], [], [], []
        off_node = off_nnz = off_map = off_row = off_rec = off_el = 0
        info = _lib.FusedInfo()
        for f in ('max_slots', 'max_nodes', 'max_rows', 'max_recs'):
            setattr(info, f, max(getattr(q.info, f) for q in plans))
        q0 = plans[0].info
        info.rec_words, info.idbits, info.offbits, info.block = (
            q0.rec_words, q0.idbits, q0.offbits, q0.block)
        offsets = []
        for q in plans:
            h = q.host
            pt = h['parts'].reshape(-1, 8).copy()
            pt[:, 0] += off_map
            pt[:, 2] += off_row
            pt[:, 5] += off_rec
            parts.append(pt.ravel())
            nodemap.append(h['nodemap'] + off_node)
            rw = h['rows'].reshape(-1, 2).copy()
            rw[:, 0] += off_nnz
            rows.append(rw.ravel())
            recs.append(h['recs'])
            rec_elem.append(h['rec_elem'] + off_el)
            offsets.append((off_node, off_nnz, off_el))
            off_node += q.nnodes
            off_nnz += q.nnz
            off_el += q.nelems
            off_map += len(h['nodemap'])
            off_row += len(h['rows']) // 2
            off_rec += len(h['rec_elem'])
        out.host = dict(parts=np.concatenate(parts), nodemap=np.concatenate(nodemap),
                        rows=np.concatenate(rows), recs=np.concatenate(recs),
                        rec_elem=np.concatenate(rec_elem))
        info.nparts = sum(q.info.nparts for q in plans)
        info.nrows = off_row
        info.nrecs = off_rec
        info.nlocal = off_map
        info.dup = float(off_rec) / max(1, off_el)
        out.info = info
        out.caps = plans[0].caps
        out.nnz, out.nnodes, out.nelems, out.dim = off_nnz, off_node, off_el, plans[0].dim
        out.offsets = offsets
        out._upload()
        return out

    def lds_bytes(self, load=False):
        return int(_lib.load().femhip_fused_lds_bytes(ctypes.byref(self.info), self.dim,
                                                      int(load)))

    def stats(self):
        i = self.info
        return dict(lds_bytes=self.lds_bytes(), parts=i.nparts, records=int(i.nrecs),
                    dup=i.dup, rec_bytes=4 * i.rec_words, max_slots=i.max_slots,
                    max_nodes=i.max_nodes, order=('colour', 'element')[self.order],
                    plan_bytes=int(sum(t.numel() * t.element_size() for t in (
                        self.parts, self.nodemap, self.rows, self.recs, self.rec_elem))))

    def check_struct(self, basis_dev, mask, nzeros):
        """The validation arguments for ``assemble(check=...)``: the
        reference tables of the basis, the plan's zero mask and count and
        the suspect queue (allocated on first use)."""
        d = basis_dev
        if self.queue is None:
            cap = max(1 << 17, self.nelems)
            self.queue = torch.empty(cap, dtype=torch.int64, device=self.parts.device)
        return _lib.FusedCheck(d.t.data_ptr(), d.dphi.data_ptr(), d.W.data_ptr(),
                               self.nelems, d.Q, 0, mask.data_ptr(), int(nzeros),
                               self.queue.data_ptr(), self.queue.numel(),
                               self.verdict.data_ptr())

    def assemble(self, name, p, out, kappa=None, load=None, groups=0, check=None):
        """out[nnz] <- CSR values of ``name`` for node coordinates ``p``
        (and load[nnodes] <- the unit-load vector when given).  With
        ``check`` (``check_struct``) the call is validated against the
        plan's pattern and the verdict {valid, mismatches, zeros, flags} is
        left in ``self.verdict`` on the device."""
        lib = _lib.load()
        assert out.numel() == self.nnz and out.dtype == torch.float64
        assert p.shape[1] == self.nnodes and p.is_contiguous()
        assert load is None or load.numel() == self.nnodes
        assert kappa is None or kappa.numel() == self.nelems
        st = self.state(name)
        _lib.check(lib.femhip_fused_assemble(
            ctypes.byref(self.dev), _lib.ptr(st), _FORM[name], p.shape[0], _lib.ptr(p),
            p.shape[1], _lib.ptr(kappa), _lib.ptr(out), _lib.ptr(load),
            None if check is None else ctypes.byref(check), int(groups),
            _lib.stream_ptr()))
        return out


# Rows plans (P2 simplices): 18,500 slots x 8 B of accumulators (148 KB of
# LDS) in one 1024-thread workgroup per CU, measured on 41^3 P2 tets against
# 16,000 / 12,000 / 9,000 / 6,000 slots with 1024 threads and 9,000 with 512
ROWS_SLOT_CAP = int(os.environ.get('FEMHIP_ROWS_SLOT_CAP', 18500))
ROWS_BLOCK = int(os.environ.get('FEMHIP_ROWS_BLOCK', 1024))
ROWS_TABLE_DOUBLES = 700   # the form tables of P2 tetrahedra (Laplace): 10*10*6 + 10*10
LDS_BYTES = 160 * 1024
ROWS_NODE_CAP = int(os.environ.get('FEMHIP_ROWS_NODE_CAP', 16384))
_ROWS_FORM = {'laplace': _lib.LAPLACE, 'mass': _lib.MASS, 'kappa_laplace': _lib.KAPPA_LAPLACE}
ROW_RECORD_WORDS = 6   # (visit << 4) | DOF, then 2 x 10 seven-bit column offsets


def rows_eligible(basis, name):
    e = basis.elem
    return (not e.vector and e.kind in ('tri', 'tet') and e.maxdeg == 2
            and e.nbs == (6 if e.kind == 'tri' else 10) and name in _ROWS_FORM)


def dof_positions(mesh, element_dofs, ndofs):
    """A point per DOF for the partition (the centroid of an element holding
    it; only the locality of the partition depends on it)."""
    t = np.asarray(mesh.t)
    c = np.asarray(mesh.p)[:, t].mean(axis=1)           # (dim, E)
    pos = np.zeros((mesh.p.shape[0], ndofs))
    ed = np.asarray(element_dofs)
    for k in range(ed.shape[0]):
        pos[:, ed[k]] = c
    return pos


class RowsFusedPlan:
    """Device arrays of a rows plan (P2 simplices): partitions of the DOF
    rows, element visits (local DOF ids) and one row record per (visit,
    owned DOF); states per form and the verdict of the last call."""

    def __init__(self, basis, indptr, indices, slot_cap=ROWS_SLOT_CAP, node_cap=ROWS_NODE_CAP):
        lib = _lib.load()
        ed = np.ascontiguousarray(basis.element_dofs, dtype=np.int32)
        N = int(basis.N)
        pos = np.ascontiguousarray(dof_positions(basis.cell_mesh, ed, N), dtype=np.float64)
        ip = np.ascontiguousarray(indptr, dtype=np.int32)
        ix = np.ascontiguousarray(indices, dtype=np.int32)
        handle = ctypes.c_void_p()
        info = _lib.FusedInfo()
        # the slot cap bounds a partition's accumulators and row count
        # together; the row tables and the form tables must fit too, so a
        # plan that overflows the LDS is rebuilt with smaller partitions
        for _ in range(8):
            _lib.check(lib.femhip_rows_fused_plan_build(
                ed.shape[0], pos.shape[0], ed.shape[1], N, ed.ctypes.data_as(ctypes.c_void_p),
                pos.ctypes.data_as(ctypes.c_void_p), ip.ctypes.data_as(ctypes.c_void_p),
                ix.ctypes.data_as(ctypes.c_void_p), int(slot_cap), int(node_cap), ROWS_BLOCK,
                ctypes.byref(handle), ctypes.byref(info)))
            if lib.femhip_rows_fused_lds_bytes(ctypes.byref(info)) + 8 * ROWS_TABLE_DOUBLES \
                    <= LDS_BYTES:
                break
            lib.femhip_fused_plan_free(handle)
            slot_cap = int(slot_cap * 0.85)
        else:
            raise ValueError("rows plan: no partition size fits the LDS")
        try:
            parts = np.empty(8 * info.nparts, np.int32)
            nodemap = np.empty(info.nlocal, np.int32)
            rows = np.empty(2 * info.nrows, np.int32)
            recs = np.empty(info.rec_words * info.nrecs, np.uint32)
            rec_elem = np.empty(info.nrecs, np.int32)
            rowrecs = np.empty(ROW_RECORD_WORDS * info.nrowrecs, np.uint32)
            arrs = (parts, nodemap, rows, recs, rec_elem)
            _lib.check(lib.femhip_fused_plan_export(
                handle, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))
            _lib.check(lib.femhip_rows_fused_plan_export(
                handle, rowrecs.ctypes.data_as(ctypes.c_void_p)))
        finally:
            lib.femhip_fused_plan_free(handle)
        dev = torch.device('cuda')
        self.info = info
        self.nnz = int(indptr[-1])
        self.parts = torch.from_numpy(parts).to(dev)
        self.nodemap = torch.from_numpy(nodemap).to(dev)
        self.rows = torch.from_numpy(rows).to(dev)
        self.recs = torch.from_numpy(recs.view(np.int32)).to(dev)
        self.rec_elem = torch.from_numpy(rec_elem).to(dev)
        self.rowrecs = torch.from_numpy(rowrecs.view(np.int32)).to(dev)
        self.dev = _lib.FusedDev(self.parts.data_ptr(), self.nodemap.data_ptr(),
                                 self.rows.data_ptr(), self.recs.data_ptr(),
                                 self.rec_elem.data_ptr(), self.info)
        self.states = {}
        self.verdict = torch.zeros(4, dtype=torch.int32, device=dev)
        self.nelems = int(ed.shape[1])
        self.mask, self.nzeros, self.queue, self.check = None, 0, None, None

    def set_pattern_zeros(self, mask, nzeros, capacity=None):
        """The validated pattern's exact-zero mask (one bit per contribution,
        ``femhip_zero_mask``) and count, and a suspect queue of ``capacity``
        entries (default: 12 per element -- the structural zeros of P2
        triangles' mass, both orders -- or twice the pattern's zeros)."""
        if capacity is None:
            capacity = max(1 << 17, 12 * self.nelems, 2 * int(nzeros))
        self.mask, self.nzeros = mask, int(nzeros)
        self.queue = torch.empty(int(capacity), dtype=torch.int64, device=self.parts.device)
        self.check = _lib.FusedCheck(None, None, None, self.nelems, 0, 0, mask.data_ptr(),
                                     self.nzeros, self.queue.data_ptr(), self.queue.numel(),
                                     self.verdict.data_ptr())

    def grow_queue(self):
        """A queue four times as long (after a call that overflowed it)."""
        self.set_pattern_zeros(self.mask, self.nzeros, 4 * self.queue.numel())

    def state(self, name):
        st = self.states.get(name)
        if st is None:
            st = torch.zeros(STATE_BYTES, dtype=torch.uint8, device=self.parts.device)
            self.states[name] = st
        return st

    def lds_bytes(self):
        return int(_lib.load().femhip_rows_fused_lds_bytes(ctypes.byref(self.info)))

    def stats(self):
        i = self.info
        return dict(lds_bytes=self.lds_bytes(), parts=i.nparts, visits=int(i.nrecs),
                    row_records=int(i.nrowrecs), dup=i.dup, max_slots=i.max_slots,
                    max_nodes=i.max_nodes)

    def assemble(self, name, d, out, kappa=None):
        """out[nnz] <- CSR values of ``name`` on the basis ``d``
        (DeviceBasis) for its current coordinates (``kappa``: per element,
        for kappa_laplace); the call is checked
        against the plan's pattern (``set_pattern_zeros``), verdict {valid,
        mismatches, zeros, flags} left in ``self.verdict`` on the device."""
        assert out.numel() == self.nnz and out.dtype == torch.float64
        assert self.check is not None, "set_pattern_zeros first"
        assert (name == 'kappa_laplace') == (kappa is not None)
        assert kappa is None or (kappa.dim() == 1 and kappa.numel() == d.E)
        metric = d.rows_metric()
        _lib.check(_lib.load().femhip_rows_fused_assemble(
            ctypes.byref(self.dev), _lib.ptr(self.rowrecs), _lib.ptr(self.state(name)),
            _ROWS_FORM[name], ctypes.byref(d.geom), ctypes.byref(d.space), _lib.ptr(metric),
            _lib.ptr(kappa), _lib.ptr(out), ctypes.byref(self.check),
            _lib.stream_ptr()))
        return out


# Laplace + mass (+ unit load) in one pass: two LDS accumulator sets, so a
# smaller partition keeps two workgroups per CU
PAIR_SLOT_CAP = int(os.environ.get('FEMHIP_PAIR_SLOT_CAP', 3072))
PAIR_NODE_CAP = int(os.environ.get('FEMHIP_PAIR_NODE_CAP', 640))


def laplace_rows(plan_rows, mass_indptr, mass_indices, lap_indptr, lap_indices):
    """Per plan row (the plan's row order: its CSR row starts in the mass
    pattern), the row's start in the Laplace pattern and the mask of the
    mass row's slots the Laplace pattern lacks (bit k: slot k)."""
    n = len(mass_indptr) - 1
    mi = np.asarray(mass_indptr, dtype=np.int64)
    li = np.asarray(lap_indptr, dtype=np.int64)
    rows_m = np.repeat(np.arange(n, dtype=np.int64), np.diff(mi))
    rows_l = np.repeat(np.arange(n, dtype=np.int64), np.diff(li))
    keys_m = rows_m * n + np.asarray(mass_indices, dtype=np.int64)
    keys_l = rows_l * n + np.asarray(lap_indices, dtype=np.int64)
    absent = ~np.isin(keys_m, keys_l)
    if not np.isin(keys_l, keys_m).all():
        raise ValueError("the Laplace pattern is not contained in the mass pattern")
    k = np.arange(len(keys_m), dtype=np.int64) - mi[rows_m]
    bits = np.zeros(n, dtype=np.int64)
    np.add.at(bits, rows_m[absent], np.left_shift(1, k[absent]))
    starts = np.asarray(plan_rows, dtype=np.int64).reshape(-1, 2)[:, 0]
    g = np.searchsorted(mi[:-1], starts, side='right') - 1
    # an empty row shares its start with the next one; take the row whose
    # start it is and that has slots
    assert np.array_equal(mi[g], starts)
    out = np.empty((len(starts), 2), dtype=np.int64)
    out[:, 0] = li[g]
    out[:, 1] = bits[g]
    return out.astype(np.uint32).view(np.int32).reshape(-1)


class LaplaceMassPass:
    """``laplace``, ``mass`` and ``unit_load`` of a P1 basis in one fused
    pass over the records and coordinates: the plan is built on the mass
    pattern, the Laplace values are written in the Laplace pattern; every
    call is checked against the Laplace pattern's zero set."""

    def __init__(self, basis, slot_cap=PAIR_SLOT_CAP, node_cap=PAIR_NODE_CAP):
        from .models import laplace, mass
        d = basis.device()
        if not (eligible(basis, 'laplace') and eligible(basis, 'mass')):
            raise NotImplementedError("the fused pair needs P1 triangles or tetrahedra")
        laplace.assemble_device(basis)       # validate both patterns (exact path)
        mass.assemble_device(basis)
        self.d = d
        self.plan_l = d.fused_last['laplace']
        self.plan_m = d.fused_last['mass']
        mi, mx = self.plan_m.indptr.cpu().numpy(), self.plan_m.indices.cpu().numpy()
        li, lx = self.plan_l.indptr.cpu().numpy(), self.plan_l.indices.cpu().numpy()
        self.fp = FusedPlan(basis.cell_mesh, mi, mx, slot_cap, node_cap)
        if self.fp.info.pruned:
            raise NotImplementedError("the mass pattern lacks an element pair")
        lr = laplace_rows(self.fp.host['rows'], mi, mx, li, lx)
        self.lrows = torch.from_numpy(lr).to(self.fp.parts.device)
        self.check = self.fp.check_struct(d, self.plan_l.mask, self.plan_l.nzeros)

    def assemble(self, laplace_out=None, mass_out=None, load_out=None, check=True):
        """(Laplace values, mass values, load vector) on the device; with
        ``check`` the verdict of the Laplace zeros is in ``self.fp.verdict``."""
        d, fp = self.d, self.fp
        dev = d.p.device
        if laplace_out is None:
            laplace_out = torch.empty(self.plan_l.nnz, dtype=torch.float64, device=dev)
        if mass_out is None:
            mass_out = torch.empty(self.plan_m.nnz, dtype=torch.float64, device=dev)
        if load_out is None:
            load_out = torch.empty(d.p.shape[1], dtype=torch.float64, device=dev)
        _lib.check(_lib.load().femhip_fused_laplace_mass(
            ctypes.byref(fp.dev), _lib.ptr(self.lrows), _lib.ptr(fp.state('pair')), d.dim,
            _lib.ptr(d.p), d.p.shape[1], _lib.ptr(laplace_out), _lib.ptr(mass_out),
            _lib.ptr(load_out), ctypes.byref(self.check) if check else None, fp.groups,
            _lib.stream_ptr()))
        return laplace_out, mass_out, load_out
