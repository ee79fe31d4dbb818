"""The fused P1 plan checked on the CPU: every index the kernel will use is
in bounds, every CSR slot is covered, colours never collide inside a
round, and replaying the kernel's accumulation order with the oracle's
local matrices reproduces the oracle's CSR matrix."""
import ctypes

import numpy as np
import pytest

import femhip
from femhip import _lib
from oracle import skfem_np as O


def _build(mesh, indptr, indices, slot_cap=6144, node_cap=1024, block=256):
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
        slot_cap, node_cap, block, ctypes.byref(h), ctypes.byref(info)))
    parts = np.empty(8 * info.nparts, np.int32)
    nodemap = np.empty(info.nlocal, np.int32)
    rows = np.empty(2 * info.nrows, np.int32)
    chunks = np.empty(2 * info.nchunks, np.int32)
    recs = np.empty(info.rec_words * info.nrecs, np.uint32)
    rec_elem = np.empty(info.nrecs, np.int32)
    arrs = (parts, nodemap, rows, chunks, recs, rec_elem)
    _lib.check(lib.femhip_fused_plan_export(h, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]))
    lib.femhip_fused_plan_free(h)
    return info, parts.reshape(-1, 8), nodemap, rows.reshape(-1, 2), chunks.reshape(-1, 2), \
        recs.reshape(-1, info.rec_words), rec_elem


def _bits(words, pos, nb):
    v = 0
    for k, w in enumerate(words):
        v |= int(w) << (32 * k)
    return (v >> pos) & ((1 << nb) - 1)


def replay(mesh, indptr, indices, Kloc, **caps):
    """Decode the plan as the kernel does; accumulate Kloc[e] (oracle local
    matrices, (E, nv, nv)) in the kernel's order; return CSR data."""
    info, parts, nodemap, rows, chunks, recs, rec_elem = _build(mesh, indptr, indices, **caps)
    nv = mesh.t.shape[0]
    idb, ofb = info.idbits, info.offbits
    nnz = int(indptr[-1])
    out = np.full(nnz, np.nan)
    written = np.zeros(nnz, int)
    lds_bytes = 8 * info.max_slots + 8 * mesh.p.shape[0] * info.max_nodes + 4 * info.max_rows
    assert lds_bytes <= 160 * 1024
    seen_rows = np.zeros(len(indptr) - 1, int)
    for (ns, L, rs, R, nslots, cs, nch, nel) in parts:
        assert L <= info.max_nodes and R <= info.max_rows and nslots <= info.max_slots
        assert L < (1 << idb)
        glob = nodemap[ns:ns + L]
        acc = np.zeros(nslots)
        rinfo = rows[rs:rs + R, 1].astype(np.uint32)
        for c in range(cs, cs + nch):
            rec0, cw = chunks[c]
            count, nrounds = cw & 0xffff, cw >> 16
            assert 0 < count <= info.block and 1 <= nrounds <= 16
            recsc = [recs[rec0 + k] for k in range(count)]
            dec = []
            for k, w in enumerate(recsc):
                lv = [_bits(w, a * idb, idb) for a in range(nv)]
                assert max(lv) < L
                offs = [_bits(w, nv * idb + f * ofb, ofb) for f in range(nv * (nv - 1))]
                rc = _bits(w, nv * idb + nv * (nv - 1) * ofb, 4)
                e = rec_elem[rec0 + k]
                assert np.array_equal(glob[lv], mesh.t[:, e])
                dec.append((lv, offs, rc, e))
            for r in range(nrounds):
                touched = set()
                for lv, offs, rc, e in dec:
                    if rc != r:
                        continue
                    for i in range(nv):
                        if lv[i] >= R:
                            continue
                        info_i = int(rinfo[lv[i]])
                        base, dg, ln = info_i & 0xffff, (info_i >> 16) & 0xff, info_i >> 24
                        assert lv[i] not in touched, "two elements of a round share a row"
                        touched.add(lv[i])
                        # a pruned entry (diagonal too) addresses the row's
                        # spare slot base + ln, which is never written out
                        assert dg <= ln
                        acc[base + dg] += Kloc[e, i, i]
                        q = 0
                        for j in range(nv):
                            if j == i:
                                continue
                            off = offs[i * (nv - 1) + q]
                            q += 1
                            assert off <= ln
                            acc[base + off] += Kloc[e, i, j]
        for r in range(R):
            gstart, info_r = rows[rs + r]
            info_r = int(np.uint32(info_r))
            base, ln = info_r & 0xffff, info_r >> 24
            assert base + ln <= nslots
            out[gstart:gstart + ln] = acc[base:base + ln]
            written[gstart:gstart + ln] += 1
            seen_rows[glob[r]] += 1
    assert np.all(written == 1), "every CSR slot is written exactly once"
    assert np.all(seen_rows == 1), "every row is owned by exactly one partition"
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
    ('tet', 7, 0.2, 'laplace', {}),
    ('tet', 6, 0.2, 'mass', {}),
    ('tet', 8, 0.2, 'laplace', {'slot_cap': 300, 'node_cap': 120}),
    ('tri', 20, 0.2, 'laplace', {'slot_cap': 100, 'node_cap': 60, 'block': 64}),
])
def test_fused_plan_replay(kind, n, a, form, caps):
    m, A, Kloc = _case(kind, n, a, form)
    out, info = replay(m, A.indptr, A.indices, Kloc, **caps)
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
                                     6144, 1024, 256, ctypes.byref(h), ctypes.byref(info))
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
        plans.append(_build(m, A.indptr, A.indices, slot_cap=700, node_cap=200, block=128))
    (i1, *a1), (i2, *a2) = plans
    assert i1.nparts == i2.nparts > 8
    for x1, x2 in zip(a1, a2):
        assert np.array_equal(x1, x2)
