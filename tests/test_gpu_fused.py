"""GPU parity of the fused P1 kernel: the second assembly on the same inputs
runs the fused kernel; it must give the first (exact-path) call's pattern
and values within 1e-12, be bitwise repeatable, and match the reference's
digests at the benchmark sizes."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import femhip
from femhip.models import laplace, mass, kappa_laplace, unit_load
from oracle import skfem_np as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
TOL = 1e-12
pytestmark = pytest.mark.gpu


def _mesh(kind, n, a):
    x = np.linspace(0, 1, n)
    m = (femhip.MeshTri.init_tensor(x, x) if kind == 'tri'
         else femhip.MeshTet.init_tensor(x, x, x))
    return m.jittered(a, 1.0 / (n - 1)) if a else m


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize('kind,n,a,form', [
    ('tri', 33, 0.0, 'laplace'), ('tri', 33, 0.2, 'laplace'), ('tri', 33, 0.2, 'mass'),
    ('tet', 9, 0.0, 'laplace'), ('tet', 12, 0.2, 'laplace'), ('tet', 9, 0.2, 'mass'),
    ('tet', 9, 0.2, 'kappa_laplace'), ('tri', 17, 0.0, 'kappa_laplace'),
])
def test_fused_matches_exact_and_oracle(kind, n, a, form):
    m = _mesh(kind, n, a)
    el = femhip.ElementTriP1() if kind == 'tri' else femhip.ElementTetP1()
    b = femhip.Basis(m, el)
    kw = {}
    if form == 'kappa_laplace':
        kw['kappa'] = np.random.default_rng(5).lognormal(0, 1, m.nelements)[:, None]
    f = {'laplace': laplace, 'mass': mass, 'kappa_laplace': kappa_laplace}[form]
    print('case', kind, n, a, form, flush=True)
    A1 = f.assemble_device(b, **kw)      # exact path (validates the pattern)
    A2 = f.assemble_device(b, **kw)      # fused kernel
    A3 = f.assemble_device(b, **kw)
    torch.cuda.synchronize()
    assert b.device().validated, "fused path was not taken"
    d1, d2, d3 = (A.data.cpu().numpy() for A in (A1, A2, A3))
    assert np.array_equal(d2, d3), "fused kernel is not bitwise repeatable"
    om = (O.MeshTri if kind == 'tri' else O.MeshTet)(m.p, m.t)
    ob = O.CellBasis(om, O.ElementTriP1() if kind == 'tri' else O.ElementTetP1())
    okw = {}
    if form == 'kappa_laplace':
        okw['kappa'] = np.repeat(kw['kappa'], ob.X.shape[1], axis=1)
    Ar = O.assemble_named(form, ob, **okw)
    for A, d in ((A1, d1), (A2, d2)):
        assert np.array_equal(A.indptr.cpu().numpy(), Ar.indptr)
        assert np.array_equal(A.indices.cpu().numpy(), Ar.indices)
        assert np.abs(d - Ar.data).max() <= TOL * np.abs(Ar.data).max()


def _check_digest(A, dig):
    assert A.nnz == dig['nnz']
    ip = A.indptr.cpu().numpy().astype(np.int32)
    ix = A.indices.cpu().numpy().astype(np.int32)
    assert _sha(ip) == dig['indptr_sha']
    assert _sha(ix) == dig['indices_sha']
    data = A.data.cpu().numpy()
    slots = np.array(dig['slots'])
    vals = np.array(dig['values'])
    scale = dig['max_abs']
    assert abs(np.abs(data).max() - scale) <= TOL * scale
    assert np.abs(data[slots] - vals).max() <= TOL * scale
    assert abs(data.sum() - dig['sum']) <= 1e-9 * max(1.0, abs(dig['sum']))


def test_tet101_digest_laplace_mass_load():
    """The benchmark mesh: 6M jittered tets, 1.03M DoF, against the
    reference's digests (pattern hashes, sampled values)."""
    dig = json.load(open(os.path.join(GOLD, 'digest_tet1_n101j.json')))
    m = _mesh('tet', 101, 0.2)
    assert _sha(m.p) == dig['p_sha'] and _sha(m.t) == dig['t_sha']
    b = femhip.Basis(m, femhip.ElementTetP1())
    for form, f in (('laplace', laplace), ('mass', mass)):
        print('form', form, flush=True)
        f.assemble_device(b)
        A = f.assemble_device(b)          # fused
        _check_digest(A, dig[form])
    u = unit_load.assemble(b)
    d = dig['unit_load']
    ref = np.array(d['values'])
    assert np.abs(u[::d['stride']][:len(ref)] - ref).max() <= TOL * np.abs(u).max()


def test_tet65_kappa_digest():
    dig = json.load(open(os.path.join(GOLD, 'digest_tet1_n65.json')))
    m = _mesh('tet', 65, 0.0)
    b = femhip.Basis(m, femhip.ElementTetP1())
    kappa = np.random.default_rng(0).lognormal(0, 1, m.nelements)
    for rep in range(2):
        A = kappa_laplace.assemble_device(b, kappa=kappa[:, None])
    _check_digest(A, dig['kappa_laplace'])
    for rep in range(2):
        A = laplace.assemble_device(b)
    _check_digest(A, dig['laplace'])


@pytest.mark.parametrize('kind,n,a,caps,block', [
    ('tet', 12, 0.2, (6144, 1024), 512), ('tet', 12, 0.2, (600, 120), 512),
    ('tri', 40, 0.2, (6144, 1024), 512), ('tri', 40, 0.0, (500, 200), 512),
    ('tet', 12, 0.2, (6144, 1024), 1024), ('tri', 40, 0.2, (700, 150), 1024),
])
def test_fixed_point_accumulation(kind, n, a, caps, block):
    """Fixed-point LDS accumulation: every form, with and without the load
    vector, persistent workgroups; values within the tolerance of the exact
    path, bitwise repeatable, and independent of what the plan assembled
    before (the exponent is re-derived from the input: a pass that ran with
    another exponent is repeated)."""
    from femhip.fused import FusedPlan
    m = _mesh(kind, n, a)
    el = femhip.ElementTriP1() if kind == 'tri' else femhip.ElementTetP1()
    b = femhip.Basis(m, el)
    d = b.device()
    rng = np.random.default_rng(11)
    k1, k2 = (rng.lognormal(0, 1, m.nelements) for _ in range(2))
    k2 *= 1e6                         # a very different scale: exponents change
    plans = {}

    def plan_for(A):
        # one plan per CSR pattern (mass keeps entries Laplace prunes)
        key = _sha(A.indices.cpu().numpy())
        if key not in plans:
            ip, ix = A.indptr.cpu().numpy(), A.indices.cpu().numpy()
            plans[key] = FusedPlan(m, ip, ix, *caps, block=block)
            assert plans[key].info.nparts > 1
        return plans[key]
    bref = unit_load.assemble(b)
    results = {}
    for step, (form, kk, load, groups) in enumerate([
            ('laplace', None, True, 0), ('mass', None, False, 5),
            ('kappa_laplace', k1, False, 0), ('kappa_laplace', k2, True, 3),
            ('kappa_laplace', k1, True, 0), ('laplace', None, False, 2),
            ('kappa_laplace', k2, False, 9)]):
        print('step', step, form, flush=True)
        kt = None if kk is None else torch.tensor(kk, device='cuda')
        kw = {} if kk is None else {'kappa': kk[:, None]}
        f = {'laplace': laplace, 'mass': mass, 'kappa_laplace': kappa_laplace}[form]
        A = f.assemble_device(b, path='generic', **kw)
        fx = plan_for(A)
        ref = A.data.cpu().numpy()
        out = torch.empty(fx.nnz, dtype=torch.float64, device='cuda')
        ld = torch.empty(b.N, dtype=torch.float64, device='cuda') if load else None
        fx.assemble(form, d.p, out, kappa=kt, load=ld, groups=groups)
        out2 = torch.empty_like(out)
        fx.assemble(form, d.p, out2, kappa=kt, groups=groups)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.abs(got - ref).max() <= TOL * np.abs(ref).max()
        assert np.array_equal(got, out2.cpu().numpy())
        if load:
            assert np.abs(ld.cpu().numpy() - bref).max() <= TOL * np.abs(bref).max()
        key = (form, None if kk is None else float(kk[0]))
        if key in results:           # same input after other inputs: same bits
            assert np.array_equal(results[key], got)
        results[key] = got


def _oracle(kind, p, t, form, kappa=None):
    om = (O.MeshTri if kind == 'tri' else O.MeshTet)(p, t)
    ob = O.CellBasis(om, O.ElementTriP1() if kind == 'tri' else O.ElementTetP1())
    if form == 'kappa_laplace':
        return O.assemble_named(form, ob, kappa=np.repeat(kappa[:, None], ob.X.shape[1], axis=1))
    return O.assemble_named(form, ob)


def _same(A, R):
    ip, ix, d = (t.cpu().numpy() for t in (A.indptr, A.indices, A.data))
    return (np.array_equal(ip, R.indptr) and np.array_equal(ix, R.indices)
            and np.abs(d - R.data).max() <= TOL * np.abs(R.data).max())


@pytest.mark.parametrize('kind,n,a', [('tet', 9, 0.2), ('tri', 33, 0.2), ('tet', 8, 0.0),
                                      ('tri', 17, 0.0)])
def test_changed_inputs_validated_in_the_fused_pass(kind, n, a):
    """New coordinates and new kappa on every call go through the fused
    kernel with the pattern check (entries near zero recomputed in the
    reference's arithmetic): every result equals the oracle's on the same
    inputs (pattern bit-exact, values within 1e-12); inputs that keep the
    plan's zero set stay on the fused path, inputs that change it (an
    unjittered mesh jittered, kappa with zeros) fall back to the exact path
    and still match."""
    m = _mesh(kind, n, a)
    el = femhip.ElementTriP1() if kind == 'tri' else femhip.ElementTetP1()
    b = femhip.Basis(m, el)
    d = b.device()
    h = 1.0 / (n - 1)
    rng = np.random.default_rng(3)
    laplace.assemble_device(b)             # exact path; the fused plan
    A = laplace.assemble_device(b)
    assert d.last_verdict is not None and d.last_verdict[0] == 1, d.last_verdict
    assert _same(A, _oracle(kind, m.p, m.t, 'laplace'))
    base = _mesh(kind, n, 0.0)
    I = base.interior_nodes()
    for step in range(3):
        p = base.p.copy()
        p[:, I] += rng.uniform(-0.2 * h, 0.2 * h, size=(p.shape[0], len(I)))
        print('coords', kind, n, a, step, flush=True)
        d.set_coordinates(torch.tensor(p, device='cuda'))
        A = laplace.assemble_device(b)
        assert _same(A, _oracle(kind, p, m.t, 'laplace')), step
        if a > 0:   # jittered -> jittered: the zero set (boundary) is the same
            assert d.last_verdict[0] == 1, d.last_verdict
    for step in range(4):
        kk = rng.lognormal(0, 1, m.nelements)
        if step == 2:
            kk[::5] = 0.0                  # zeros prune more entries
        print('kappa', kind, n, a, step, flush=True)
        A = kappa_laplace.assemble_device(b, kappa=torch.tensor(kk, device='cuda')[:, None])
        assert _same(A, _oracle(kind, d.p.cpu().numpy(), m.t, 'kappa_laplace', kk)), step


def test_unfusable_mesh_falls_back_to_exact_path():
    """A node with 40 neighbours (a CSR row of 41 entries, more than the
    fused records can encode): the fused plan is refused and every call
    stays on the exact path, matching the oracle."""
    K = 40
    ang = 2 * np.pi * np.arange(K) / K
    p = np.vstack([np.r_[0.0, np.cos(ang)], np.r_[0.0, np.sin(ang)]])
    t = np.array([[0] * K, 1 + np.arange(K), 1 + (np.arange(K) + 1) % K])
    m = femhip.MeshTri(p, t)
    b = femhip.Basis(m, femhip.ElementTriP1())
    R = O.bilinear_assemble(O.form_laplace, O.CellBasis(O.MeshTri(m.p, m.t), O.ElementTriP1()))
    for _ in range(3):
        A = laplace.assemble(b)
        assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
        assert np.abs(A.data - R.data).max() <= TOL * np.abs(R.data).max()
    d = b.device()
    assert d.unfusable and not d.validated
