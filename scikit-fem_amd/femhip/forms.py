"""``BilinearForm`` / ``LinearForm`` / ``asm`` with scikit-fem's call
signatures (``skfem/assembly/form/form.py:28-121``,
``bilinear_form.py:130-148``, ``linear_form.py:18-49``,
``skfem/assembly/__init__.py:69-97``).

A form is accelerated when it is one of the built-in models of
``femhip.models`` (``laplace``, ``mass``, ``unit_load``, ``kappa_laplace``,
``linear_elasticity(Lambda, Mu)``), or a user form that the tracer in
``femhip.trace`` recognises as one of them, or as a coefficient times
``u * v`` (``coef_mass``) or times ``v`` (``coef_load``), times a constant.
On a ``FacetBasis`` the accelerated forms are ``mass``, ``unit_load``,
``coef_mass`` and ``coef_load``.  Anything else raises
``NotImplementedError``: there is no CPU fallback.
"""
import numbers

import numpy as np
import torch

from .trace import identify


class Form:

    form = None

    def __init__(self, form=None, dtype=np.float64, nthreads=0, **params):
        self.form = form.form if isinstance(form, Form) else form
        self.dtype = dtype
        self.nthreads = nthreads
        self.params = params
        self._spec = getattr(form, '_spec', None)

    def __call__(self, *args):
        if self.form is None:  # used as a decorator with arguments
            return type(self)(form=args[0], dtype=self.dtype,
                              nthreads=self.nthreads, **self.params)
        raise NotImplementedError("direct evaluation of forms is not supported")

    # identification of the built-in kernel ---------------------------------
    def spec(self):
        """(kernel name, constant factor, parameters) of this form."""
        if self._spec is None:
            self._spec = identify(self.form, self.nargs)
        return self._spec

    def _check_dtype(self):
        if np.dtype(self.dtype) != np.float64:
            raise NotImplementedError("only float64 assembly is accelerated")


_UPLOAD_CACHE = {}


def _kappa_tensor(w, basis, name='kappa'):
    """Coefficient array of ``w[name]`` as a device tensor: (E,) per element
    or (E, Q) per quadrature point (the reference accepts an (E, Q) array or
    a DiscreteField, ``form.py:92-121``); a 1-D array of length N is a DOF
    vector, interpolated at the quadrature points of ``basis``
    (``form.py:106-108``).  An (E, Q) array with identical columns is treated
    as per-element.

    Uploads of host arrays are cached by a 128-bit content digest (the
    interpolated DOF vectors on the basis itself, with its coordinate
    version).  Plan reuse does not rely on these caches: the assembly
    compares the coefficient with the one its plan was validated for."""
    if name not in w:
        raise ValueError("Attribute '%s' not found in 'w'." % name)
    k = w[name]
    if hasattr(k, 'value') and not isinstance(k, np.ndarray):
        k = k.value
    E, Q = basis.nelems, basis.W.shape[0]
    if isinstance(k, torch.Tensor):
        kt = k.detach().to(dtype=torch.float64, device='cuda')
    else:
        import hashlib
        k = np.asarray(k, dtype=np.float64)
        if k.ndim == 1:
            # a DOF vector: interpolated at the quadrature points
            # (form.py:106-108), cached per basis and coordinate version
            if k.shape[0] != basis.N:
                raise ValueError("Input array has wrong size.")
            dev = basis.device()
            key = (hashlib.blake2b(k.tobytes(), digest_size=16).hexdigest(), dev.p_version)
            cache = getattr(dev, 'coef_cache', None)
            kt = None if cache is None else cache.get(key)
            if kt is None:
                kt = basis.interpolate_device(k, grad=False).value.reshape(E, Q).contiguous()
                dev.coef_cache = {key: kt}
            return kt
        if k.ndim == 2 and k.shape == (E, Q) and Q > 1 and np.all(k == k[:, :1]):
            k = k[:, :1]
        k = np.ascontiguousarray(k)
        key = (k.shape, hashlib.blake2b(k.tobytes(), digest_size=16).hexdigest())
        kt = _UPLOAD_CACHE.get(key)
        if kt is None:
            kt = torch.from_numpy(k).to('cuda')
            _UPLOAD_CACHE.clear()
            _UPLOAD_CACHE[key] = kt
    if kt.dim() == 2 and kt.shape == (E, 1):
        return kt.reshape(E).contiguous()
    if kt.dim() == 2 and kt.shape == (E, Q):
        return kt.reshape(E, Q).contiguous()
    if kt.dim() == 1 and kt.shape[0] == E:
        return kt.contiguous()
    raise ValueError("Quadrature mismatch: '%s' should have shape (E, Q) = "
                     "(%d, %d) or (E, 1)" % (name, E, Q))


class BilinearForm(Form):
    """Bilinear form ``a(u, v)``; ``assemble`` returns a SciPy CSR matrix
    with the reference's canonical pattern."""

    nargs = 3

    def assemble_device(self, ubasis, vbasis=None, path='auto', **kwargs):
        """Assemble and keep the result in HBM (``DeviceCSR``).  ``path``:
        'auto' (fused kernel when available), 'generic' (exact path)."""
        self._check_dtype()
        if vbasis is not None and vbasis is not ubasis:
            return self._assemble_mixed(ubasis, vbasis, **kwargs)
        name, scale, params = self.spec()
        if name == 'vector_laplace':
            A = _vector_laplace(ubasis, path)
            if scale != 1.0:
                A.data = A.data * scale
            return A
        dev = ubasis.device()
        if name in ('kappa_laplace', 'coef_mass'):
            kappa = _kappa_tensor({**self.params, **kwargs}, ubasis,
                                  params.get('coef', 'kappa'))
            A = dev.assemble_bilinear(name, kappa=kappa, path=path)
        elif name == 'linear_elasticity':
            A = dev.assemble_bilinear(name, lam=params['Lambda'], mu=params['Mu'], path=path)
        else:
            A = dev.assemble_bilinear(name, path=path)
        if scale != 1.0:
            A.data = A.data * scale
        return A

    def _assemble_mixed(self, ubasis, vbasis, path='auto', **kwargs):
        """Trial functions from ``ubasis``, test functions from ``vbasis``
        (``bilinear_form.py:58-128``): rows are the test basis's DOFs, the
        shape is (vbasis.N, ubasis.N), the weights and coefficients are the
        trial basis's.  Both bases are cell bases on the same elements with
        the same quadrature points, both scalar or both vector."""
        if ubasis.X.shape[-1] != vbasis.X.shape[-1]:
            raise ValueError("Quadrature mismatch: trial and test functions "
                             "should have same number of integration points.")
        for b in (ubasis, vbasis):
            if getattr(b, 'find', None) is not None:
                raise NotImplementedError("distinct trial and test bases: cell bases only")
        same_cells = (vbasis.mesh is ubasis.mesh or (
            np.array_equal(vbasis.mesh.p, ubasis.mesh.p)
            and np.array_equal(vbasis.mesh.t, ubasis.mesh.t)))
        same_tind = ((ubasis.tind is None and vbasis.tind is None) or (
            ubasis.tind is not None and vbasis.tind is not None
            and np.array_equal(ubasis.tind, vbasis.tind)))
        if not (same_cells and same_tind):
            raise NotImplementedError("distinct trial and test bases on different elements")
        if not (np.array_equal(ubasis.X, vbasis.X) and np.array_equal(ubasis.W, vbasis.W)):
            raise NotImplementedError("distinct trial and test bases: different quadrature points")
        if bool(ubasis.elem.vector) != bool(vbasis.elem.vector):
            raise NotImplementedError("a vector trial basis with a scalar test basis")
        name, scale, params = self.spec()
        if name not in ('laplace', 'mass', 'kappa_laplace', 'coef_mass', 'linear_elasticity'):
            raise NotImplementedError("distinct trial and test bases: form %r" % (name,))
        kappa = None
        if name in ('kappa_laplace', 'coef_mass'):
            kappa = _kappa_tensor({**self.params, **kwargs}, ubasis,
                                  params.get('coef', 'kappa'))
        A = ubasis.device().assemble_mixed(vbasis.device(), name, kappa=kappa,
                                           lam=params.get('Lambda', 0.0),
                                           mu=params.get('Mu', 0.0))
        if scale != 1.0:
            A.data = A.data * scale
        return A

    def assemble(self, ubasis, vbasis=None, **kwargs):
        return self.assemble_device(ubasis, vbasis, **kwargs).to_scipy()

    def elemental(self, ubasis, **kwargs):
        """Local element matrices (E, Nbfun, Nbfun) as ``COOData.tolocal``
        returns them (``coo_data.py:45-66``)."""
        name, scale, params = self.spec()
        dev = ubasis.device()
        kappa = None
        if name in ('kappa_laplace', 'coef_mass'):
            kappa = _kappa_tensor({**self.params, **kwargs}, ubasis,
                                  params.get('coef', 'kappa'))
        c = dev.local_bilinear(name, kappa=kappa, lam=params.get('Lambda', 0.0),
                               mu=params.get('Mu', 0.0))
        nb = ubasis.Nbfun
        loc = c.reshape(nb, nb, ubasis.nelems).permute(2, 1, 0)  # (e, i, j)
        loc = loc.cpu().numpy() * scale
        return loc


def _vector_laplace(basis, path):
    """``vector_laplace`` (``models/poisson.py:12-14``) on ElementVector(e):
    ``laplace`` on e (fused for P1 simplices) expanded to the components --
    the reference's ``ddot(grad u, grad v)`` adds the scalar dot product's
    terms in the same order for equal components and exact zeros otherwise,
    so the values and the pruned pattern are the scalar ones, repeated."""
    import ctypes
    from . import _lib
    from .device import DeviceCSR
    if not basis.elem.vector or getattr(basis, 'find', None) is not None:
        raise NotImplementedError("vector_laplace on a CellBasis of ElementVector")
    sb = getattr(basis, '_scalar_companion', None)
    if sb is None:
        sb = basis.with_element(basis.elem.elem)
        basis._scalar_companion = sb
    S = sb.device().assemble_bilinear('laplace', path=path)
    dim = basis.elem.dim
    n, nnz = S.shape[0], S.nnz
    dev = S.data.device
    indptr = torch.empty(dim * n + 1, dtype=torch.int32, device=dev)
    indices = torch.empty(max(dim * nnz, 1), dtype=torch.int32, device=dev)[:dim * nnz]
    data = torch.empty(max(dim * nnz, 1), dtype=torch.float64, device=dev)[:dim * nnz]
    _lib.check(_lib.load().femhip_csr_expand_components(
        _lib.ptr(S.indptr), _lib.ptr(S.indices), _lib.ptr(S.data), n, nnz, dim,
        _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(data), _lib.stream_ptr()))
    return DeviceCSR(indptr, indices, data, (dim * n, dim * n))


class LinearForm(Form):
    """Linear form ``l(v)``; ``assemble`` returns the load vector."""

    nargs = 2

    def assemble_device(self, basis, vbasis=None, **kwargs):
        self._check_dtype()
        if vbasis is not None:
            raise ValueError("LinearForm takes one basis")
        name, scale, params = self.spec()
        if name == 'coef_load':
            kappa = _kappa_tensor({**self.params, **kwargs}, basis, params['coef'])
            b = basis.device().assemble_linear('coef_load', kappa=kappa)
        elif name == 'unit_load':
            b = basis.device().assemble_linear('unit_load')
        else:
            raise NotImplementedError("linear forms: unit_load and coef_load")
        if scale != 1.0:
            b = b * scale
        return b

    def assemble(self, basis, vbasis=None, **kwargs):
        from .device import to_host
        return to_host(self.assemble_device(basis, vbasis, **kwargs))


def asm(form, *bases, **kwargs):
    """``skfem.assembly.asm`` (``assembly/__init__.py:63-97``): a form, or a
    plain function of 2 (linear) or 3 (bilinear) arguments, on one basis or
    on a list of bases whose results are summed.

    The reference sums a list by concatenating the COO entries of every
    basis and converting once (``coo_data.py:28-36,78-90``): exact-zero
    entries are dropped, every other slot stays, also a sum that cancels to
    0.0.  Here each basis is assembled on the device and the CSR results are
    concatenated as COO entries and converted once, so the pattern is the
    reference's; values agree to rounding (a slot's contributions are added
    per basis first).  A trial and a test argument, each a basis or a list,
    assemble every (trial, test) combination, as the reference does."""
    if not isinstance(form, Form):
        if not callable(form):
            raise ValueError("unknown form type")
        import inspect
        n = len(inspect.signature(form).parameters)
        if n not in (2, 3):
            raise NotImplementedError("asm: linear and bilinear forms only")
        form = BilinearForm(form) if n == 3 else LinearForm(form)
    if not bases or len(bases) > 2:
        raise ValueError("asm takes one basis, or a trial and a test basis")
    lists = [list(b) if isinstance(b, (list, tuple)) else [b] for b in bases]
    if any(not b for b in lists):
        raise ValueError("asm: empty list of bases")
    if isinstance(form, LinearForm) and len(lists) != 1:
        raise ValueError("LinearForm takes one basis")
    # every combination of the arguments, summed (assembly/__init__.py:91-94)
    import itertools
    combos = list(itertools.product(*lists))
    if len(combos) == 1:
        return form.assemble(*combos[0], **kwargs)
    parts = [form.assemble(*c, **kwargs) for c in combos]
    if isinstance(form, LinearForm):
        out = parts[0].copy()
        for v in parts[1:]:
            out += v
        return out
    return _coo_sum(parts)


def _coo_sum(mats):
    """CSR matrices summed as one COO list converted to CSR (explicit zeros
    of cancelled sums are kept, as the reference's conversion keeps them)."""
    from scipy.sparse import coo_matrix
    shape = tuple(max(m.shape[i] for m in mats) for i in range(2))
    rows, cols, data = [], [], []
    for m in mats:
        c = m.tocoo()
        rows.append(c.row)
        cols.append(c.col)
        data.append(c.data)
    K = coo_matrix((np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))),
                   shape=shape)
    return K.tocsr()


def is_number(x):
    return isinstance(x, numbers.Number)
