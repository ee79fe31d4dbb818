"""The layer-sweep plan of the fused P1 kernel checked on the CPU: replaying
the kernel's decoding, level ring buffers and colour rounds with the
oracle's local matrices reproduces the oracle's CSR matrix and load vector;
every index the kernel uses is in bounds and no two elements of a round
touch the same row."""
import numpy as np
import pytest

from femhip.fused import sweep_plan_arrays
from oracle import skfem_np as O


def _bits(words, pos, nb):
    v = 0
    for k, w in enumerate(words):
        v |= int(w) << (32 * k)
    return (v >> pos) & ((1 << nb) - 1)


def replay(mesh, A, Kloc, vols, block_rows, chunk):
    info, h = sweep_plan_arrays(mesh, A.indptr, A.indices, block_rows, chunk)
    nv = mesh.t.shape[0]
    dim = mesh.p.shape[0]
    ofb = info.offbits
    sent = (1 << ofb) - 1
    rw = info.rec_words
    blocks = h['blocks'].reshape(-1, 4)
    levels = h['levels'].reshape(-1, 6)
    rows = h['rows'].reshape(-1, 2)
    chunks = h['chunks'].reshape(-1, 2)
    recs = h['recs'].reshape(-1, rw)
    out = np.full(A.nnz, np.nan)
    written = np.zeros(A.nnz, int)
    load = np.full(mesh.p.shape[1], np.nan)
    MS = info.max_level_slots
    for ls, nl, _, _ in blocks:
        Rr = [0, 0]
        acc = [np.zeros(MS), np.zeros(MS)]
        bacc = [np.zeros(info.max_level_rows), np.zeros(info.max_level_rows)]
        win = [None, None]
        rinfo = [None, None]

        def prep(l):
            node_off, nn, row_off, nr, ns, _ = levels[ls + l]
            assert nn <= info.max_level_nodes and nr <= info.max_level_rows and ns <= MS
            b = l & 1
            acc[b][:] = 0.0
            bacc[b][:] = 0.0
            win[b] = h['nodemap'][node_off:node_off + nn]
            rinfo[b] = rows[row_off:row_off + nr, 1].astype(np.uint32)
            Rr[b] = nr
        prep(0)
        for l in range(nl):
            if l + 1 < nl:
                prep(l + 1)
            else:
                Rr[(l + 1) & 1] = 0
            c0, c1 = levels[ls + l][5], levels[ls + l + 1][5]
            for c in range(c0, c1):
                rec0, cw = chunks[c]
                count, nrounds = cw & 0xffff, cw >> 16
                assert 0 < count <= chunk and 1 <= nrounds <= 16
                dec = []
                for k in range(count):
                    w = recs[rec0 + k]
                    e = h['rec_elem'][rec0 + k]
                    ids = [_bits(w, 11 * a, 11) for a in range(nv)]
                    par = [i >> 10 for i in ids]
                    idx = [i & 1023 for i in ids]
                    for a in range(nv):
                        assert idx[a] < len(win[par[a]])
                        assert win[par[a]][idx[a]] == mesh.t[a, e]
                    offs = [_bits(w, nv * 11 + f * ofb, ofb) for f in range(nv * (nv - 1))]
                    rc = _bits(w, nv * 11 + nv * (nv - 1) * ofb, 4)
                    dec.append((par, idx, offs, rc, e))
                for r in range(nrounds):
                    touched = set()
                    for par, idx, offs, rc, e in dec:
                        if rc != r:
                            continue
                        for i in range(nv):
                            b = par[i]
                            if idx[i] >= Rr[b]:
                                continue
                            assert (b, idx[i]) not in touched
                            touched.add((b, idx[i]))
                            bacc[b][idx[i]] += vols[e]
                            inf = int(rinfo[b][idx[i]])
                            base, dg, ln = inf & 0xffff, (inf >> 16) & 0xff, inf >> 24
                            if dg != 0xff:
                                acc[b][base + dg] += Kloc[e, i, i]
                            q = 0
                            for j in range(nv):
                                if j == i:
                                    continue
                                off = offs[i * (nv - 1) + q]
                                q += 1
                                if off != sent:
                                    assert off < ln
                                    acc[b][base + off] += Kloc[e, i, j]
            node_off, nn, row_off, nr, ns, _ = levels[ls + l]
            b = l & 1
            for r in range(nr):
                gstart, inf = rows[row_off + r]
                inf = int(np.uint32(inf))
                base, ln = inf & 0xffff, inf >> 24
                assert base + ln <= ns
                out[gstart:gstart + ln] = acc[b][base:base + ln]
                written[gstart:gstart + ln] += 1
                load[h['nodemap'][node_off + r]] = bacc[b][r]
    assert np.all(written == 1)
    return out, load, info


@pytest.mark.parametrize('kind,n,a,br,ck', [
    ('tet', 7, 0.2, 60, 512), ('tet', 6, 0.0, 4096, 512), ('tet', 8, 0.2, 100, 64),
    ('tri', 17, 0.2, 40, 512), ('tri', 13, 0.0, 1000, 1024),
])
def test_sweep_plan_replay(kind, n, a, br, ck):
    x = np.linspace(0, 1, n)
    if kind == 'tri':
        m, e, fac = O.MeshTri.init_tensor(x, x), O.ElementTriP1(), 1.0 / 6.0
    else:
        m, e, fac = O.MeshTet.init_tensor(x, x, x), O.ElementTetP1(), 1.0 / 24.0
    if a:
        m = O.jitter(m, a, 1.0 / (n - 1))
    b = O.CellBasis(m, e)
    idx, data, shape = O.bilinear_coo(O.form_laplace, b)
    A = O.coo_to_csr(idx, data, shape)
    nb = m.t.shape[0]
    Kloc = data.reshape(nb, nb, -1).transpose(2, 1, 0)
    vols = np.abs(b.mapping.detA) * fac
    out, load, info = replay(m, A, Kloc, vols, br, ck)
    assert np.abs(out - A.data).max() <= 1e-13 * np.abs(A.data).max()
    bref = O.linear_assemble(O.form_unit_load, b)
    assert np.abs(load - bref).max() <= 1e-13 * np.abs(bref).max()
    assert info.dup >= 1.0
