"""Batched assembly on the GPU: many kappa samples on one mesh (shared
pattern), and many meshes of uneven size in one (graph-captured) launch."""
import numpy as np
import pytest
import torch

import femhip
from femhip.batch import KappaSamples, MeshBatch
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _embed(indptr, indices, Aref):
    """Reference CSR values placed into a superset pattern (0 elsewhere)."""
    n = len(indptr) - 1
    out = np.zeros(indptr[-1])
    for r in range(n):
        cols = indices[indptr[r]:indptr[r + 1]]
        rc = Aref.indices[Aref.indptr[r]:Aref.indptr[r + 1]]
        pos = np.searchsorted(cols, rc)
        assert np.array_equal(cols[pos], rc), "reference slot missing from the shared pattern"
        out[indptr[r] + pos] = Aref.data[Aref.indptr[r]:Aref.indptr[r + 1]]
    return out


@pytest.mark.parametrize('el,n', [('tri2', 9), ('tri1', 17), ('tet1', 5)])
def test_kappa_samples_match_per_sample_reference(el, n):
    x = np.linspace(0, 1, n)
    if el.startswith('tri'):
        m = femhip.MeshTri.init_tensor(x, x)
        e = femhip.ElementTriP2() if el == 'tri2' else femhip.ElementTriP1()
        om = O.MeshTri(m.p, m.t)
        oe = O.ElementTriP2() if el == 'tri2' else O.ElementTriP1()
    else:
        m = femhip.MeshTet.init_tensor(x, x, x)
        e, om, oe = femhip.ElementTetP1(), O.MeshTet(m.p, m.t), O.ElementTetP1()
    b = femhip.Basis(m, e)
    S = 5
    kap = np.random.default_rng(0).lognormal(0, 0.5, (S, m.nelements))
    ks = KappaSamples(b)
    vals = ks.assemble(torch.tensor(kap, device='cuda')).cpu().numpy()
    ip, ix = ks.indptr.cpu().numpy(), ks.indices.cpu().numpy()
    ob = O.CellBasis(om, oe)
    for s in range(S):
        Ar = O.assemble_named('kappa_laplace', ob,
                              kappa=np.repeat(kap[s][:, None], ob.X.shape[1], axis=1))
        ref = _embed(ip, ix, Ar)
        assert np.abs(vals[s] - ref).max() <= TOL * np.abs(ref).max(), s
    # repeatable
    vals2 = ks.assemble(torch.tensor(kap, device='cuda')).cpu().numpy()
    assert np.array_equal(vals, vals2)


def _meshes(k=6, seed=0):
    rng = np.random.default_rng(seed)
    ns = rng.integers(5, 40, k)
    out = []
    for i, n in enumerate(ns):
        x = np.linspace(0, 1, int(n))
        m = femhip.MeshTri.init_tensor(x, x)
        if i % 2:
            m = m.jittered(0.2, 1.0 / (n - 1), seed=i)
        out.append(m)
    return out


@pytest.mark.parametrize('groups', [None, 8])
def test_mesh_batch_one_launch(groups):
    ms = _meshes()
    bases = [femhip.Basis(m, femhip.ElementTriP1()) for m in ms]
    mb = MeshBatch(bases, groups=groups)
    mb.assemble()
    torch.cuda.synchronize()
    for (ip, ix, data, load), m in zip(mb.split(), ms):
        ob = O.CellBasis(O.MeshTri(m.p, m.t), O.ElementTriP1())
        Ar = O.bilinear_assemble(O.form_laplace, ob)
        br = O.linear_assemble(O.form_unit_load, ob)
        assert np.array_equal(ip.cpu().numpy(), Ar.indptr)
        assert np.array_equal(ix.cpu().numpy(), Ar.indices)
        d = data.cpu().numpy()
        assert np.abs(d - Ar.data).max() <= TOL * np.abs(Ar.data).max()
        assert np.abs(load.cpu().numpy() - br).max() <= TOL * np.abs(br).max()


def test_mesh_batch_graph_replays():
    """Every replay of the captured launch equals the eager result, with
    eager work and synchronisations between the replays."""
    ms = _meshes(k=5, seed=3)
    mb = MeshBatch([femhip.Basis(m, femhip.ElementTriP1()) for m in ms], groups=16)
    d0, l0 = (t.clone() for t in mb.assemble())
    torch.cuda.synchronize()
    mb.capture()
    for rep in range(4):
        mb.data.zero_()
        mb.load.fill_(-1.0)
        torch.cuda.synchronize()
        d, l = mb.replay()
        torch.cuda.synchronize()
        assert torch.equal(d, d0) and torch.equal(l, l0), rep
        _ = (torch.ones(1000, device='cuda') * rep).sum().item()   # eager work


def test_kappa_samples_graph_replays():
    """The captured κ-sample launch: new samples copied into the graph's
    input buffer before each replay; every replay equals the eager launch on
    the same samples (eager work and synchronisations between replays)."""
    x = np.linspace(0, 1, 13)
    m = femhip.MeshTri.init_tensor(x, x)
    b = femhip.Basis(m, femhip.ElementTriP2())
    ks = KappaSamples(b)
    S = 4
    rng = np.random.default_rng(11)
    kap = torch.tensor(rng.lognormal(0, 0.5, (S, m.nelements)), device='cuda')
    out = torch.empty((S, ks.nnz), dtype=torch.float64, device='cuda')
    ks.capture(kap, out)
    for r in range(4):
        print('replay', r, flush=True)
        new = torch.tensor(rng.lognormal(0, 0.5, (S, m.nelements)), device='cuda')
        kap.copy_(new)
        ks.replay()
        torch.cuda.synchronize()
        ref = ks.assemble(new.clone())
        torch.cuda.synchronize()
        assert torch.equal(out, ref), r
