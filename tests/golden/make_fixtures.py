"""Write the golden fixtures of the assembly path from the reference itself.

Run once in the development container (the reference is importable there
only; it never travels anywhere else):

    python tests/golden/make_fixtures.py [--big]

What is written (all under tests/golden/):

* ``small_*.npz``  -- full inputs (p, t) and full outputs (CSR indptr,
  indices, data, or the load vector) of small meshes, one per element/form
  case.  These pin the exact-zero-pruned CSR pattern, which no reference
  test pins.
* ``config1_tri129.npz`` -- the full CSR stiffness matrix and load vector of
  the 129 x 129 unit-square P1 case.
* ``digest_*.json`` (``--big``) -- SHA-256 of indptr/indices, nnz, max|A|,
  sum(data) and 4096 seeded (slot, value) samples of the large cases, plus
  the SHA-256 of the jittered node array the outputs were computed on.
* ``facet_*.npz`` -- boundary forms on a ``FacetBasis``: the facets and
  their owning elements, the facet quadrature, dx, the outward normals, the
  CSR facet mass matrix and the facet load vector (default boundary facets,
  and a set of interior facets seen from either side); the facet points and
  mesh parameter, a Robin matrix ``alpha * u * v`` and a Neumann vector
  ``g * v`` with data computed at the facet points, a Neumann vector whose
  data is a DOF vector (interpolated by the reference), and that
  interpolation (value and gradient).
* ``coef_*.npz`` -- the same coefficient forms on a ``CellBasis`` (the
  quadrature points and mesh parameter, ``alpha * u * v``, ``g * v``).
* ``grad_kappa65.npz`` (``--big``) -- the reference gradient of
  L = 1/2 u^T A(kappa, X) u - f^T u on the 65^3 tet mesh: dL/dkappa_e from
  single-element assemblies (unit coefficient) and dL/dX from central
  differences of the reference assembly on node patches.

Only data is stored: inputs and expected outputs.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, '/root/reference')

import skfem  # noqa: E402
from skfem import (MeshTri, MeshTet, MeshHex, ElementTriP1, ElementTriP2,  # noqa: E402
                   ElementTetP1, ElementTetP2, ElementHex1, ElementHex2,
                   ElementVector, Basis, CellBasis, BilinearForm, LinearForm)
from skfem.helpers import dot, grad  # noqa: E402
from skfem.models.poisson import laplace, mass, unit_load, vector_laplace  # noqa: E402
from skfem.models.elasticity import linear_elasticity, lame_parameters  # noqa: E402


@BilinearForm
def kappa_laplace(u, v, w):
    return dot(w['kappa'] * grad(u), grad(v))


@BilinearForm
def robin(u, v, w):
    return w['alpha'] * u * v


@LinearForm
def neumann(v, w):
    return w['g'] * v


def coef_data(x):
    """Robin and Neumann data at points x (dim, n, Q)."""
    alpha = 1.0 + x[0] ** 2 + 0.5 * x[1]
    g = np.sin(3.0 * x[0]) + x[1] * x[-1]
    return alpha, g


def jitter(mesh, a, h, seed=0):
    rng = np.random.default_rng(seed)
    I = mesh.interior_nodes()
    p = mesh.p.copy()
    p[:, I] += rng.uniform(-a * h, a * h, size=(p.shape[0], len(I)))
    return type(mesh)(p, mesh.t)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def mk_mesh(kind, n, a=0.0):
    x = np.linspace(0, 1, n)
    h = 1.0 / (n - 1)
    m = {'tri': lambda: MeshTri.init_tensor(x, x),
         'tet': lambda: MeshTet.init_tensor(x, x, x),
         'hex': lambda: MeshHex.init_tensor(x, x, x)}[kind]()
    if a > 0:
        m = jitter(m, a, h)
    return m


ELEMS = {
    'tri1': ElementTriP1, 'tri2': ElementTriP2, 'tet1': ElementTetP1,
    'tet2': ElementTetP2, 'hex1': ElementHex1, 'hex2': ElementHex2,
}


def make_basis(m, elem, vector=False, intorder=None):
    e = ELEMS[elem]()
    if vector:
        e = ElementVector(e)
    return Basis(m, e, intorder=intorder)


def assemble(form, basis, kappa_e=None):
    lam, mu = lame_parameters(1.0, 0.3)
    if form == 'laplace':
        return laplace.assemble(basis)
    if form == 'mass':
        return mass.assemble(basis)
    if form == 'unit_load':
        return unit_load.assemble(basis)
    if form == 'kappa_laplace':
        kq = np.repeat(kappa_e[:, None], basis.X.shape[1], axis=1)
        return kappa_laplace.assemble(basis, kappa=kq)
    if form == 'linear_elasticity':
        return linear_elasticity(lam, mu).assemble(basis)
    if form == 'vector_laplace':
        return vector_laplace.assemble(basis)
    raise KeyError(form)


SMALL = [
    # name, mesh kind, n, jitter a, element, vector, intorder, forms
    ('tri1_n9', 'tri', 9, 0.0, 'tri1', False, None,
     ['laplace', 'mass', 'unit_load', 'kappa_laplace']),
    ('tri1_n9j', 'tri', 9, 0.2, 'tri1', False, None, ['laplace', 'mass']),
    ('tri2_n17', 'tri', 17, 0.0, 'tri2', False, None,
     ['laplace', 'mass', 'unit_load', 'kappa_laplace']),
    ('tri2_n9j', 'tri', 9, 0.2, 'tri2', False, None, ['laplace']),
    ('tet1_n6', 'tet', 6, 0.0, 'tet1', False, None,
     ['laplace', 'mass', 'unit_load', 'kappa_laplace']),
    ('tet1_n7j', 'tet', 7, 0.2, 'tet1', False, None,
     ['laplace', 'mass', 'unit_load', 'kappa_laplace']),
    ('tet2_n5', 'tet', 5, 0.0, 'tet2', False, None, ['laplace', 'mass']),
    ('tet2_n5j', 'tet', 5, 0.2, 'tet2', False, None,
     ['laplace', 'mass', 'unit_load']),
    ('hex1_n5', 'hex', 5, 0.0, 'hex1', False, None, ['laplace', 'mass']),
    ('hex1_n5j', 'hex', 5, 0.15, 'hex1', False, None,
     ['laplace', 'mass', 'unit_load']),
    ('hex2_n4j', 'hex', 4, 0.15, 'hex2', False, 5, ['laplace', 'mass']),
    ('hex2_n3', 'hex', 3, 0.0, 'hex2', False, 5, ['laplace']),
    ('vhex2_n4j', 'hex', 4, 0.15, 'hex2', True, 5, ['linear_elasticity']),
    ('vhex2_n3', 'hex', 3, 0.0, 'hex2', True, 5, ['linear_elasticity']),
    ('vtet1_n5j', 'tet', 5, 0.2, 'tet1', True, None, ['linear_elasticity', 'vector_laplace']),
    ('vtri2_n5j', 'tri', 5, 0.2, 'tri2', True, None, ['linear_elasticity', 'vector_laplace']),
]


def write_small():
    for name, kind, n, a, elem, vec, io, forms in SMALL:
        m = mk_mesh(kind, n, a)
        b = make_basis(m, elem, vec, io)
        out = {'p': m.p, 't': m.t, 'element_dofs': b.element_dofs,
               'N': np.array(b.N), 'nq': np.array(b.X.shape[1])}
        rng = np.random.default_rng(7)
        kappa_e = rng.lognormal(0.0, 1.0, m.t.shape[1])
        out['kappa_e'] = kappa_e
        for f in forms:
            r = assemble(f, b, kappa_e)
            if f == 'unit_load':
                out[f + '_b'] = r
            else:
                r = r.tocsr()
                out[f + '_indptr'] = r.indptr
                out[f + '_indices'] = r.indices
                out[f + '_data'] = r.data
        meta = dict(kind=kind, n=n, jitter=a, element=elem, vector=vec,
                    intorder=-1 if io is None else io, forms=forms)
        np.savez_compressed(os.path.join(HERE, 'small_%s.npz' % name),
                            meta=json.dumps(meta), **out)
        print('wrote', name, {f: out.get(f + '_data', np.empty(0)).size
                              for f in forms})


FACET = [
    # name, mesh kind, n, jitter a, element, facet selection, side
    ('tri1_n9j', 'tri', 9, 0.2, 'tri1', None, 0),
    ('tri2_n9j', 'tri', 9, 0.2, 'tri2', None, 0),
    ('tri1_n9j_mid1', 'tri', 9, 0.2, 'tri1', 'mid', 1),
    ('tet1_n6j', 'tet', 6, 0.2, 'tet1', None, 0),
    ('tet2_n5j', 'tet', 5, 0.2, 'tet2', None, 0),
    ('tet2_n5j_mid0', 'tet', 5, 0.2, 'tet2', 'mid', 0),
    ('tet1_n6', 'tet', 6, 0.0, 'tet1', None, 0),
]


def write_facet():
    from skfem import FacetBasis
    for name, kind, n, a, elem, sel, side in FACET:
        m = mk_mesh(kind, n, a)
        facets = None
        if sel == 'mid':
            # interior facets on the grid plane x = 0.5 (n odd), selected on
            # the undistorted mesh (same t, hence same facet numbering)
            m0 = mk_mesh(kind, n)
            facets = m0.facets_satisfying(lambda x: np.abs(x[0] - 0.5) < 1e-9)
        fb = FacetBasis(m, ELEMS[elem](), facets=facets, side=side)
        A = mass.assemble(fb).tocsr()
        f = unit_load.assemble(fb)
        x = np.asarray(fb.global_coordinates())
        alpha, g = coef_data(x)
        R = robin.assemble(fb, alpha=alpha).tocsr()
        wdof = 1.0 + 0.5 * np.cos(np.arange(fb.N, dtype=np.float64))
        ip = fb.interpolate(wdof)
        out = dict(p=m.p, t=m.t, find=fb.find, tind=fb.tind,
                   element_dofs=fb.element_dofs, N=np.array(fb.N), X=fb.X, W=fb.W,
                   dx=fb.dx, normals=fb.normals.value,
                   mass_indptr=A.indptr, mass_indices=A.indices, mass_data=A.data,
                   unit_load_b=f, x=x, h=np.asarray(fb.mesh_parameters()),
                   alpha=alpha, g=g, robin_indptr=R.indptr, robin_indices=R.indices,
                   robin_data=R.data, neumann_b=neumann.assemble(fb, g=g),
                   wdof=wdof, neumann_dof_b=neumann.assemble(fb, g=wdof),
                   interp_value=np.asarray(ip), interp_grad=ip.grad)
        meta = dict(kind=kind, n=n, jitter=a, element=elem,
                    facets='boundary' if sel is None else sel, side=side)
        np.savez_compressed(os.path.join(HERE, 'facet_%s.npz' % name),
                            meta=json.dumps(meta), **out)
        print('wrote facet', name, len(fb.find), 'facets, nnz', A.nnz)


COEF = [
    # name, mesh kind, n, jitter a, element
    ('tri2_n9j', 'tri', 9, 0.2, 'tri2'),
    ('tet1_n7j', 'tet', 7, 0.2, 'tet1'),
    ('tet2_n5j', 'tet', 5, 0.2, 'tet2'),
    ('hex1_n5j', 'hex', 5, 0.15, 'hex1'),
]


def write_coef():
    for name, kind, n, a, elem in COEF:
        m = mk_mesh(kind, n, a)
        b = make_basis(m, elem)
        x = np.asarray(b.global_coordinates())
        alpha, g = coef_data(x)
        R = robin.assemble(b, alpha=alpha).tocsr()
        wdof = 1.0 + 0.5 * np.cos(np.arange(b.N, dtype=np.float64))
        out = dict(p=m.p, t=m.t, N=np.array(b.N), x=x, h=np.asarray(b.mesh_parameters()),
                   alpha=alpha, g=g, robin_indptr=R.indptr, robin_indices=R.indices,
                   robin_data=R.data, neumann_b=neumann.assemble(b, g=g),
                   wdof=wdof, neumann_dof_b=neumann.assemble(b, g=wdof))
        meta = dict(kind=kind, n=n, jitter=a, element=elem)
        np.savez_compressed(os.path.join(HERE, 'coef_%s.npz' % name),
                            meta=json.dumps(meta), **out)
        print('wrote coef', name, R.nnz)


def write_config1():
    m = mk_mesh('tri', 129)
    b = Basis(m, ElementTriP1())
    A = laplace.assemble(b)
    f = unit_load.assemble(b)
    D = m.boundary_nodes()
    np.savez_compressed(os.path.join(HERE, 'config1_tri129.npz'),
                        indptr=A.indptr, indices=A.indices, data=A.data,
                        b=f, D=D, p_sha=sha(m.p))
    print('config1 nnz', A.nnz)


def digest(A, nsamp=4096, seed=0):
    A = A.tocsr()
    rng = np.random.default_rng(seed)
    slots = np.sort(rng.choice(A.nnz, size=min(nsamp, A.nnz), replace=False))
    return dict(shape=list(A.shape), nnz=int(A.nnz),
                indptr_sha=sha(A.indptr.astype(np.int32)),
                indices_sha=sha(A.indices.astype(np.int32)),
                max_abs=float(np.abs(A.data).max()),
                sum=float(A.data.sum()),
                slots=slots.tolist(), values=A.data[slots].tolist())


def write_digest(name, kind, n, a, elem, vec, io, forms):
    t0 = time.time()
    m = mk_mesh(kind, n, a)
    b = make_basis(m, elem, vec, io)
    rec = dict(kind=kind, n=n, jitter=a, element=elem, vector=vec,
               intorder=-1 if io is None else io, p_sha=sha(m.p),
               t_sha=sha(m.t), N=int(b.N), nelems=int(m.t.shape[1]))
    for f in forms:
        t1 = time.time()
        if f == 'kappa_laplace':
            kappa_e = np.random.default_rng(0).lognormal(0, 1, m.t.shape[1])
            r = assemble(f, b, kappa_e)
        else:
            r = assemble(f, b)
        dt = time.time() - t1
        if f == 'unit_load':
            rec[f] = dict(sha=sha(r), sum=float(r.sum()),
                          values=r[::max(1, len(r) // 4096)].tolist(),
                          stride=max(1, len(r) // 4096), seconds=dt)
        else:
            rec[f] = digest(r)
            rec[f]['seconds'] = dt
        print(name, f, 'done in %.1fs' % dt, flush=True)
    with open(os.path.join(HERE, 'digest_%s.json' % name), 'w') as fh:
        json.dump(rec, fh)
    print('wrote digest', name, '%.1fs' % (time.time() - t0), flush=True)


def write_grad_kappa(n=65, nsel=64, step=1e-6):
    """Reference gradient of L = 1/2 u^T A u - f^T u, A = kappa-Laplace."""
    x = np.linspace(0, 1, n)
    m = MeshTet.init_tensor(x, x, x)
    E = m.t.shape[1]
    N = m.p.shape[1]
    kappa = np.random.default_rng(0).lognormal(0.0, 1.0, E)
    u = np.random.default_rng(1).standard_normal(N)
    f = np.random.default_rng(2).standard_normal(N)
    rng = np.random.default_rng(3)
    els = np.sort(rng.choice(E, size=nsel, replace=False))
    nodes = np.sort(rng.choice(m.interior_nodes(), size=nsel, replace=False))
    e = ElementTetP1()
    # dL/dkappa_e = 1/2 u^T K_e(kappa=1) u from a single-element assembly
    dk = np.zeros(nsel)
    for k, el in enumerate(els):
        bb = CellBasis(m, e, elements=np.array([el]))
        Ke = laplace.assemble(bb)
        dk[k] = 0.5 * u @ (Ke @ u)
    # dL/dX by central differences of the patch assembly
    t = m.t
    dX = np.zeros((nsel, 3))
    for k, nd in enumerate(nodes):
        patch = np.nonzero((t == nd).any(axis=0))[0]
        for c in range(3):
            vals = []
            for s in (+1, -1):
                p = m.p.copy()
                p[c, nd] += s * step
                mm = MeshTet(p, t)
                bb = CellBasis(mm, e, elements=patch)
                kq = np.repeat(kappa[patch][:, None], bb.X.shape[1], axis=1)
                A = kappa_laplace.assemble(bb, kappa=kq)
                vals.append(0.5 * u @ (A @ u))
            dX[k, c] = (vals[0] - vals[1]) / (2 * step)
    # the loss value itself on the full mesh
    b = Basis(m, e)
    kq = np.repeat(kappa[:, None], b.X.shape[1], axis=1)
    A = kappa_laplace.assemble(b, kappa=kq)
    L = 0.5 * u @ (A @ u) - f @ u
    np.savez_compressed(os.path.join(HERE, 'grad_kappa65.npz'),
                        elements=els, nodes=nodes, dkappa=dk, dX=dX,
                        L=np.array(L), step=np.array(step), n=np.array(n))
    print('wrote grad_kappa65', L)


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--big', action='store_true')
    ap.add_argument('--only', default=None)
    args = ap.parse_args()
    print('reference version', skfem.__version__)
    if args.only in (None, 'small'):
        write_small()
        write_config1()
    if args.only in (None, 'facet'):
        write_facet()
    if args.only in (None, 'coef'):
        write_coef()
    if args.big:
        big = {
            'tet1_n101j': ('tet', 101, 0.2, 'tet1', False, None,
                           ['laplace', 'mass', 'unit_load']),
            'tet2_n41j': ('tet', 41, 0.2, 'tet2', False, 4, ['laplace']),
            'tri2_n257': ('tri', 257, 0.0, 'tri2', False, None,
                          ['laplace', 'kappa_laplace']),
            'tet1_n65': ('tet', 65, 0.0, 'tet1', False, None,
                         ['laplace', 'kappa_laplace']),
            'vhex2_n11j': ('hex', 11, 0.15, 'hex2', True, 5,
                           ['linear_elasticity']),
            'vhex2_n21j': ('hex', 21, 0.15, 'hex2', True, 5,
                           ['linear_elasticity']),
        }
        for k, v in big.items():
            if args.only in (None, k):
                write_digest(k, *v)
        if args.only in (None, 'grad'):
            write_grad_kappa()
