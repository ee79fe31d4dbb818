"""Every benchmark configuration pinned at the size the reference fixtures
cover (tests/golden/digest_*.json, written from the reference itself:
SHA-256 of indptr/indices, nnz, max|A|, sum and 4096 sampled slots):

* P2 tetrahedra 41^3 jittered, Laplace at intorder 4 -- the first call
  (exact path, pattern built) and a later call (validated pattern);
* P2 triangles 257^2, Laplace and kappa-Laplace with kappa ~ lognormal(0, 1)
  (seed 0) -- and that kappa as sample 0 of a 256-sample KappaSamples batch
  (its sampled slots, through the export to the reference's pruned CSR);
* the 64 P1 triangle meshes of the batching benchmark
  (n_i = default_rng(0).integers(17, 258, 64)) in one MeshBatch launch,
  every mesh against the oracle;
* vector Q2 hexahedral elasticity on the 20^3 jittered mesh (MFMA path)."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import femhip
from femhip.models import laplace, kappa_laplace, linear_elasticity, lame_parameters
from oracle import skfem_np as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
TOL = 1e-12
pytestmark = pytest.mark.gpu


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _check(A, dig):
    assert A.nnz == dig['nnz']
    assert _sha(A.indptr.cpu().numpy().astype(np.int32)) == dig['indptr_sha']
    assert _sha(A.indices.cpu().numpy().astype(np.int32)) == dig['indices_sha']
    data = A.data.cpu().numpy()
    scale = dig['max_abs']
    assert abs(np.abs(data).max() - scale) <= TOL * scale
    assert np.abs(data[np.array(dig['slots'])] - np.array(dig['values'])).max() <= TOL * scale
    assert abs(data.sum() - dig['sum']) <= 1e-9 * max(1.0, abs(dig['sum']))


def test_p2_tet_41_digest_first_and_later_calls():
    dig = json.load(open(os.path.join(GOLD, 'digest_tet2_n41j.json')))
    x = np.linspace(0, 1, 41)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1 / 40)
    assert _sha(m.p) == dig['p_sha'] and _sha(m.t) == dig['t_sha']
    b = femhip.Basis(m, femhip.ElementTetP2(), intorder=4)
    assert b.N == dig['N']
    print('first call', flush=True)
    _check(laplace.assemble_device(b), dig['laplace'])
    print('later call', flush=True)
    _check(laplace.assemble_device(b), dig['laplace'])
    d = b.device()
    assert 'laplace' in d.rows_last and d.last_verdict[0] == 1   # fused rows kernel


def test_p2_tri_257_digests_and_kappa_batch():
    from femhip.batch import KappaSamples
    dig = json.load(open(os.path.join(GOLD, 'digest_tri2_n257.json')))
    x = np.linspace(0, 1, 257)
    m = femhip.MeshTri.init_tensor(x, x)
    assert _sha(m.p) == dig['p_sha'] and _sha(m.t) == dig['t_sha']
    b = femhip.Basis(m, femhip.ElementTriP2())
    print('laplace', flush=True)
    _check(laplace.assemble_device(b), dig['laplace'])
    kappa = np.random.default_rng(0).lognormal(0, 1, m.nelements)
    print('kappa_laplace', flush=True)
    _check(kappa_laplace.assemble_device(b, kappa=kappa[:, None]), dig['kappa_laplace'])
    # the same kappa as sample 0 of a 256-sample batch (other samples:
    # lognormal(0, 0.5) as in the batching benchmark)
    S = 256
    kap = np.random.default_rng(1).lognormal(0, 0.5, (S, m.nelements))
    kap[0] = kappa
    ks = KappaSamples(b)
    print('batch', flush=True)
    vals = ks.assemble(torch.tensor(kap, device='cuda'))
    A0 = ks.export(0, torch.tensor(kappa, device='cuda'), vals)
    _check(A0, dig['kappa_laplace'])
    A7 = ks.export(7, torch.tensor(kap[7], device='cuda'), vals)
    R7 = kappa_laplace.assemble_device(b, kappa=kap[7][:, None])
    assert torch.equal(A7.indptr, R7.indptr) and torch.equal(A7.indices, R7.indices)
    assert (A7.data - R7.data).abs().max() <= TOL * R7.data.abs().max()


def test_mesh_batch_64_benchmark_meshes():
    from femhip.batch import MeshBatch
    ns = np.random.default_rng(0).integers(17, 258, 64)
    bases = []
    for n in ns:
        x = np.linspace(0, 1, int(n))
        bases.append(femhip.Basis(femhip.MeshTri.init_tensor(x, x), femhip.ElementTriP1()))
    print('64 meshes', int(sum(b.nelems for b in bases)), 'triangles', flush=True)
    mb = MeshBatch(bases, groups=512)
    mb.capture()
    mb.replay()
    torch.cuda.synchronize()
    for k, ((ip, ix, data, load), bb) in enumerate(zip(mb.split(), bases)):
        m = bb.mesh
        ob = O.CellBasis(O.MeshTri(m.p, m.t), O.ElementTriP1())
        Ar = O.bilinear_assemble(O.form_laplace, ob)
        br = O.linear_assemble(O.form_unit_load, ob)
        assert np.array_equal(ip.cpu().numpy(), Ar.indptr), k
        assert np.array_equal(ix.cpu().numpy(), Ar.indices), k
        assert np.abs(data.cpu().numpy() - Ar.data).max() <= TOL * np.abs(Ar.data).max(), k
        assert np.abs(load.cpu().numpy() - br).max() <= TOL * np.abs(br).max(), k


def test_elasticity_20_digest():
    fn = os.path.join(GOLD, 'digest_vhex2_n21j.json')
    dig = json.load(open(fn))
    x = np.linspace(0, 1, 21)
    m = femhip.MeshHex.init_tensor(x, x, x).jittered(0.15, 1 / 20)
    assert _sha(m.p) == dig['p_sha']
    b = femhip.Basis(m, femhip.ElementVector(femhip.ElementHex2()), intorder=5)
    lam, mu = lame_parameters(1.0, 0.3)
    form = linear_elasticity(lam, mu)
    print('first call', flush=True)
    _check(form.assemble_device(b), dig['linear_elasticity'])
    print('MFMA call', flush=True)
    A = form.assemble_device(b)
    assert b.device().validated
    _check(A, dig['linear_elasticity'])
