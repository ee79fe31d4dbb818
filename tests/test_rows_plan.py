"""The rows plan of P2 simplices (fused_plan.cpp, rows mode), built by the
C ABI on the host, and a NumPy replay of the fused rows kernel's schedule on
it: every (element, DOF) pair is owned by exactly one partition, the
records name the element's DOFs, and summing the oracle's local entries
in that schedule -- a pair of two owned DOFs computed once by the lower
local index and added to both rows, a pair with a DOF owned elsewhere added
to this row only, exact zeros skipped -- gives the oracle's CSR matrix,
every pattern slot receiving a nonzero contribution."""
import ctypes

import numpy as np
import pytest

import femhip
from femhip import _lib
from femhip.fused import dof_positions
from oracle import skfem_np as O


def _rows_plan(basis, indptr, indices, slot_cap, node_cap):
    lib = _lib.load()
    ed = np.ascontiguousarray(basis.element_dofs, dtype=np.int32)
    N = int(basis.N)
    pos = np.ascontiguousarray(dof_positions(basis.mesh, ed, N), dtype=np.float64)
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    h = ctypes.c_void_p()
    info = _lib.FusedInfo()
    _lib.check(lib.femhip_rows_fused_plan_build(
        ed.shape[0], pos.shape[0], ed.shape[1], N, ed.ctypes.data_as(ctypes.c_void_p),
        pos.ctypes.data_as(ctypes.c_void_p), ip.ctypes.data_as(ctypes.c_void_p),
        ix.ctypes.data_as(ctypes.c_void_p), slot_cap, node_cap, 256, ctypes.byref(h),
        ctypes.byref(info)))
    try:
        out = dict(parts=np.empty(8 * info.nparts, np.int32),
                   nodemap=np.empty(info.nlocal, np.int32),
                   rows=np.empty(2 * info.nrows, np.int32),
                   recs=np.empty(info.rec_words * info.nrecs, np.uint32),
                   rec_elem=np.empty(info.nrecs, np.int32))
        _lib.check(lib.femhip_fused_plan_export(
            h, *[out[k].ctypes.data_as(ctypes.c_void_p)
                 for k in ('parts', 'nodemap', 'rows', 'recs', 'rec_elem')]))
        out['rowrecs'] = np.empty(6 * info.nrowrecs, np.uint32)
        _lib.check(lib.femhip_rows_fused_plan_export(
            h, out['rowrecs'].ctypes.data_as(ctypes.c_void_p)))
    finally:
        lib.femhip_fused_plan_free(h)
    return info, out


@pytest.mark.parametrize('kind,n,a,form,caps', [
    ('tri', 9, 0.2, 'laplace', (3072, 1024)), ('tri', 9, 0.0, 'laplace', (200, 64)),
    ('tet', 4, 0.2, 'laplace', (400, 100)), ('tet', 4, 0.0, 'mass', (400, 100)),
    ('tet', 5, 0.0, 'laplace', (3072, 1024)),
])
def test_rows_plan_replay_gives_the_oracle_matrix(kind, n, a, form, caps):
    x = np.linspace(0, 1, n)
    if kind == 'tri':
        m = femhip.MeshTri.init_tensor(x, x)
        el, om_t, oel = femhip.ElementTriP2(), O.MeshTri, O.ElementTriP2()
    else:
        m = femhip.MeshTet.init_tensor(x, x, x)
        el, om_t, oel = femhip.ElementTetP2(), O.MeshTet, O.ElementTetP2()
    if a:
        m = m.jittered(a, 1.0 / (n - 1))
    b = femhip.Basis(m, el)
    ob = O.CellBasis(om_t(m.p, m.t), oel)
    f = {'laplace': O.form_laplace, 'mass': O.form_mass}[form]
    coo, data, shape = O.bilinear_coo(f, ob)
    Aref = O.coo_to_csr(coo, data, shape)
    nb, E = ob.Nbfun, ob.nelems
    K = data.reshape(nb, nb, E)          # K[j, i, e]: row dof i, column dof j
    info, P = _rows_plan(b, Aref.indptr, Aref.indices, *caps)
    ed = np.asarray(b.element_dofs)
    assert np.array_equal(ed, ob.element_dofs)
    RW = info.rec_words
    out = np.zeros(Aref.nnz)
    got = np.zeros(Aref.nnz, dtype=bool)
    seen = np.zeros((nb, E), dtype=np.int64)

    def slot(g, c):
        s0, s1 = Aref.indptr[g], Aref.indptr[g + 1]
        k = s0 + np.searchsorted(Aref.indices[s0:s1], c)
        assert k < s1 and Aref.indices[k] == c, "nonzero entry outside the pattern"
        return k

    for p in range(info.nparts):
        node_start, Ln, row_start, R, nslots, rr0, nrr, v0 = P['parts'][8 * p:8 * p + 8]
        assert Ln <= info.max_nodes and R <= info.max_rows and nslots <= info.max_slots
        gid = P['nodemap'][node_start:node_start + Ln]
        rows = P['rows'][2 * row_start:2 * (row_start + R)].reshape(R, 2)
        for r in range(R):   # row table: CSR start, LDS base, length
            g = gid[r]
            base, ln = rows[r, 1] & 0xffff, (rows[r, 1] >> 24) & 0x7f
            assert rows[r, 0] == Aref.indptr[g] and ln == Aref.indptr[g + 1] - Aref.indptr[g]
            assert base + ln <= nslots
        for k in range(nrr):
            rec = P['rowrecs'][6 * (rr0 + k):6 * (rr0 + k + 1)].astype(np.int64)
            v, i = int(rec[0]) >> 4, int(rec[0]) & 15
            field = [int(rec[1 + (f >> 2)] >> (7 * (f & 3))) & 127 for f in range(2 * nb)]
            ri = sum(int(rec[1 + q] >> 28) << (4 * q) for q in range(4))
            w = P['recs'][v * RW:(v + 1) * RW]
            lid = [int((w[j >> 1] >> ((j & 1) * 16)) & 0xffff) for j in range(nb)]
            e = P['rec_elem'][v]
            assert w[RW - 1] == e                 # the element id closes the record
            assert all(lid[j] < Ln and gid[lid[j]] == ed[j, e] for j in range(nb))
            assert lid[i] < R and ri == lid[i]
            seen[i, e] += 1
            for j in range(nb):
                own = lid[j] < R
                if own and j < i:
                    continue
                r = K[j, i, e]
                if r == 0.0:
                    continue
                s = slot(ed[i, e], ed[j, e])      # raises for a pair outside the pattern
                assert field[j] != 127 and rows[ri, 0] + field[j] == s
                out[s] += r
                got[s] = True
                if own and j != i:
                    assert K[i, j, e] == r       # symmetric bit for bit
                    s = slot(ed[j, e], ed[i, e])
                    assert field[nb + j] != 127 and rows[lid[j], 0] + field[nb + j] == s
                    out[s] += r
                    got[s] = True
    assert (seen == 1).all()
    assert got.all()
    assert np.abs(out - Aref.data).max() <= 1e-12 * np.abs(Aref.data).max()


def test_rows_plan_rejects_p1():
    x = np.linspace(0, 1, 5)
    m = femhip.MeshTri.init_tensor(x, x)
    b = femhip.Basis(m, femhip.ElementTriP1())
    ip = np.arange(b.N + 1, dtype=np.int32)
    with pytest.raises(NotImplementedError):
        _rows_plan(b, ip, np.arange(b.N, dtype=np.int32), 3072, 1024)
