"""The fused P1 plan checked on the CPU: every index the kernel will use is
in bounds, every CSR slot is covered exactly once, every row is owned by
one partition, and replaying the kernel's accumulation (off-diagonal pairs
added to both owned rows, the Laplace diagonal as the negated sum of the
row's other slots and its spare slot, the mass diagonal added) with the
oracle's local matrices reproduces the oracle's CSR matrix."""
import ctypes

import numpy as np
import pytest

import femhip
from femhip import _lib
from oracle import skfem_np as O


def _build(mesh, indptr, indices, slot_cap=6144, node_cap=1024, block=512, order=1):
    lib = _lib.load()
    t = np.ascontiguousarray(mesh.t, np.int32)
    p = np.ascontiguousarray(mesh.p, np.float64)
    ip = np.ascontiguousarray(indptr, np.int32)
    ix = np.ascontiguousarray(indices, np.int32)
    h = ctypes.c_void_p()
    info = _lib.FusedInfo()
    _lib.check(lib.femhip_fused_plan_build(
        t.shape[0], p.shape[0], t.shape[1], p.shape[1],
        t.ctypes.data_as(ctypes.c_void_p), p.ctypes.data_as(ctypes.c_void_p),
        ip.ctypes.data_as(ctypes.c_void_p), ix.ctypes.data_as(ctypes.c_void_p),
        slot_cap, node_cap, block, order, ctypes.byref(h), ctypes.byref(info)))
    parts = np.empty(8 * info.nparts, np.int32)
    nodemap = np.empty(info.nlocal, np.int32)
    rows = np.empty(2 * info.nrows, np.int32)
    recs = np.empty(info.rec_words * info.nrecs, np.uint32)
    rec_elem = np.empty(info.nrecs, np.int32)
    arrs = (parts, nodemap, rows, recs, rec_elem)
    _lib.check(lib.femhip_fused_plan_export(h, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))
    lib.femhip_fused_plan_free(h)
    return info, parts.reshape(-1, 8), nodemap, rows.reshape(-1, 2), \
        recs.reshape(-1, info.rec_words), rec_elem


def _bits(words, pos, nb):
    v = 0
    for k, w in enumerate(words):
        v |= int(w) << (32 * k)
    return (v >> pos) & ((1 << nb) - 1)


def replay(mesh, indptr, indices, Kloc, rowsum, **caps):
    """Decode the plan as the kernel does; accumulate Kloc[e] (oracle local
    matrices, (E, nv, nv), [e, row i, column j]) as the kernel does; return
    the CSR data."""
    info, parts, nodemap, rows, recs, rec_elem = _build(mesh, indptr, indices, **caps)
    nv = mesh.t.shape[0]
    dim = mesh.p.shape[0]
    idb, ofb = info.idbits, info.offbits
    nnz = int(indptr[-1])
    out = np.full(nnz, np.nan)
    written = np.zeros(nnz, int)
    lib = _lib.load()
    assert lib.femhip_fused_lds_bytes(ctypes.byref(info), dim, 1) <= 160 * 1024
    assert (info.rec_words * 32 >= nv * idb + nv * (nv - 1) * ofb)
    seen_rows = np.zeros(len(indptr) - 1, int)
    seen_elem = np.zeros(mesh.t.shape[1], int)
    for (ns, L, rs, R, nslots, r0, nrec, _) in parts:
        assert L <= info.max_nodes and R <= info.max_rows and nslots <= info.max_slots
        assert nrec <= info.max_recs and L < (1 << idb)
        glob = nodemap[ns:ns + L]
        assert np.all(np.diff(glob[:R]) > 0) and np.all(np.diff(glob[R:]) > 0)
        acc = np.zeros(nslots)
        rinfo = rows[rs:rs + R, 1].view(np.uint32)
        # LDS rows back to back: base_{r+1} = base_r + len_r + spare_r
        base_next = 0
        for r in range(R):
            ri = int(rinfo[r])
            assert ri & 0xffff == base_next
            base_next += ((ri >> 24) & 0x7f) + (ri >> 31)
        assert base_next == nslots
        for k in range(r0, r0 + nrec):
            w = recs[k]
            lv = [_bits(w, a * idb, idb) for a in range(nv)]
            assert max(lv) < L
            offs = [_bits(w, nv * idb + f * ofb, ofb) for f in range(nv * (nv - 1))]
            e = rec_elem[k]
            assert np.array_equal(glob[lv], mesh.t[:, e])
            assert min(lv) < R, "a visited element has an owned vertex"
            seen_elem[e] += sum(1 for v in lv if v < R) > 0
            rbase = {}
            for i in range(nv):
                if lv[i] < R:
                    ri = int(rinfo[lv[i]])
                    rbase[i] = (ri & 0xffff, (ri >> 16) & 0xff, (ri >> 24) & 0x7f)
            for i in range(nv):
                for j in range(i + 1, nv):
                    if i in rbase:
                        off = offs[i * (nv - 1) + j - 1]
                        assert off <= rbase[i][2]
                        acc[rbase[i][0] + off] += Kloc[e, i, j]
                    if j in rbase:
                        off = offs[j * (nv - 1) + i]
                        assert off <= rbase[j][2]
                        acc[rbase[j][0] + off] += Kloc[e, j, i]
            if not rowsum:
                for i, (b, dg, ln) in rbase.items():
                    assert dg <= ln
                    acc[b + dg] += Kloc[e, i, i]
        for r in range(R):
            gstart, ri = rows[rs + r]
            ri = int(np.uint32(ri))
            b, dg, ln, sp = ri & 0xffff, (ri >> 16) & 0xff, (ri >> 24) & 0x7f, ri >> 31
            assert b + ln + sp <= nslots
            vals = acc[b:b + ln].copy()
            if rowsum and dg < ln:
                vals[dg] = -(acc[b:b + ln].sum() - acc[b + dg] + (acc[b + ln] if sp else 0.0))
            out[gstart:gstart + ln] = vals
            written[gstart:gstart + ln] += 1
            seen_rows[glob[r]] += 1
    assert np.all(written == 1), "every CSR slot is written exactly once"
    assert np.all(seen_rows == 1), "every row is owned by exactly one partition"
    assert np.all(seen_elem >= 1), "every element is visited"
    return out, info


def _case(kind, n, a, form):
    x = np.linspace(0, 1, n)
    if kind == 'tri':
        m = O.MeshTri.init_tensor(x, x)
        e = O.ElementTriP1()
    else:
        m = O.MeshTet.init_tensor(x, x, x)
        e = O.ElementTetP1()
    if a:
        m = O.jitter(m, a, 1.0 / (n - 1))
    b = O.CellBasis(m, e)
    f = {'laplace': O.form_laplace, 'mass': O.form_mass}[form]
    idx, data, shape = O.bilinear_coo(f, b)
    A = O.coo_to_csr(idx, data, shape)
    nb = m.t.shape[0]
    Kloc = data.reshape(nb, nb, -1).transpose(2, 1, 0)  # (e, i=row, j=col)
    return m, A, Kloc


@pytest.mark.parametrize('kind,n,a,form,caps', [
    ('tri', 17, 0.0, 'laplace', {}),
    ('tri', 17, 0.2, 'mass', {}),
    ('tet', 7, 0.0, 'laplace', {}),
    ('tet', 7, 0.2, 'laplace', {'order': 0}),
    ('tet', 6, 0.2, 'mass', {}),
    ('tet', 8, 0.2, 'laplace', {'slot_cap': 300, 'node_cap': 120}),
    ('tri', 20, 0.2, 'laplace', {'slot_cap': 100, 'node_cap': 60, 'order': 0}),
])
def test_fused_plan_replay(kind, n, a, form, caps):
    m, A, Kloc = _case(kind, n, a, form)
    out, info = replay(m, A.indptr, A.indices, Kloc, form == 'laplace', **caps)
    assert np.abs(out - A.data).max() <= 1e-13 * np.abs(A.data).max()
    assert info.dup >= 1.0


def test_fused_plan_rejects_long_rows():
    lib = _lib.load()
    m = femhip.MeshTri()
    indptr = np.array([0, 40, 40, 40, 40], np.int32)
    indices = np.zeros(40, np.int32)
    h = ctypes.c_void_p()
    info = _lib.FusedInfo()
    t = np.ascontiguousarray(m.t)
    p = np.ascontiguousarray(m.p)
    rc = lib.femhip_fused_plan_build(3, 2, t.shape[1], p.shape[1],
                                     t.ctypes.data_as(ctypes.c_void_p),
                                     p.ctypes.data_as(ctypes.c_void_p),
                                     indptr.ctypes.data_as(ctypes.c_void_p),
                                     indices.ctypes.data_as(ctypes.c_void_p),
                                     6144, 1024, 512, 1, ctypes.byref(h), ctypes.byref(info))
    assert rc == _lib.EUNSUPPORTED


def test_plan_independent_of_thread_count(monkeypatch):
    """The blocks are built by a pool of host threads and concatenated in
    block order: one thread and five threads give the same plan."""
    x = np.linspace(0, 1, 13)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1 / 12)
    om = O.MeshTet(m.p, m.t)
    A = O.bilinear_assemble(O.form_laplace, O.CellBasis(om, O.ElementTetP1()))
    plans = []
    for nth in ('1', '5'):
        monkeypatch.setenv('FEMHIP_PLAN_THREADS', nth)
        plans.append(_build(m, A.indptr, A.indices, slot_cap=700, node_cap=200, order=0))
    (i1, *a1), (i2, *a2) = plans
    assert i1.nparts == i2.nparts > 8
    for x1, x2 in zip(a1, a2):
        assert np.array_equal(x1, x2)
