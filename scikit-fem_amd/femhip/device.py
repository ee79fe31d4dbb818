"""Device side of a ``CellBasis``: uploaded mesh arrays, the symbolic plans
and the launch sequence of the generic assembly path.

Generic path (every element / form):

1. ``femhip_geometry`` -- mapping (invDF, |detDF|) in the reference's
   operation order, once per basis;
2. ``femhip_local_bilinear`` -- all local entries, bit-identical to the
   reference's NumPy values, in its COO order;
3. symbolic plan -- CSR pattern of the nonzero entries + per-slot
   contribution ranges, built once and revalidated on every call by an
   exact comparison of the exact-zero positions with the plan's bit mask;
4. ``femhip_segment_sum`` -- deterministic per-slot sums.

The fused P1 path (``femhip.fused``) replaces 2-4 for P1 triangles and
tetrahedra once the plan exists.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("femhip needs a ROCm GPU (torch.cuda is unavailable)")
    return torch.device('cuda')


def check_host_inputs(p, t, dofs, N):
    """Validation before anything is uploaded or indexed on the device:
    finite node coordinates, connectivity inside the node array and DOF
    indices inside [0, N) (a DOF outside the matrix would corrupt the
    symbolic phase).  Raises ValueError."""
    if not np.all(np.isfinite(p)):
        raise ValueError("mesh coordinates are not finite")
    if t.size and (t.min() < 0 or t.max() >= p.shape[1]):
        raise ValueError("element connectivity index out of range [0, %d)" % p.shape[1])
    if dofs.size and (dofs.min() < 0 or dofs.max() >= N):
        raise ValueError("element DOF index out of range [0, %d): min %d, max %d"
                         % (N, int(dofs.min()), int(dofs.max())))


def to_dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(_dev())


_KIND = {'tri': _lib.TRI, 'tet': _lib.TET, 'hex': _lib.HEX}
_FORM = {'laplace': _lib.LAPLACE, 'mass': _lib.MASS,
         'kappa_laplace': _lib.KAPPA_LAPLACE, 'coef_mass': _lib.COEF_MASS,
         'coef_load': _lib.COEF_LOAD,
         'linear_elasticity': _lib.ELASTICITY, 'unit_load': _lib.UNIT_LOAD}


_PINNED = {}


def to_host(t):
    """A fresh NumPy copy of a device tensor, through a pinned staging
    buffer kept per dtype (a pageable copy runs at a fraction of the link's
    rate)."""
    n = t.numel()
    buf = _PINNED.get(t.dtype)
    if buf is None or buf.numel() < n:
        buf = torch.empty(max(n, 1), dtype=t.dtype, pin_memory=True)
        _PINNED[t.dtype] = buf
    view = buf[:n]
    view.copy_(t.reshape(-1))           # ordered after the producing kernels
    out = np.empty(tuple(t.shape), dtype=view.numpy().dtype)
    np.copyto(out.reshape(-1), view.numpy())
    return out


class DeviceCSR:
    """A CSR matrix whose arrays live in HBM (int32 indptr/indices, float64
    data).  ``to_scipy`` copies it to the host as ``csr_matrix``.  Matrices
    assembled on one pattern share ``plan``, which keeps a host copy of the
    pattern after the first transfer (each matrix gets its own copy of it)."""

    def __init__(self, indptr, indices, data, shape, plan=None):
        self.indptr, self.indices, self.data, self.shape = indptr, indices, data, shape
        self.plan = plan

    @property
    def nnz(self):
        return int(self.indices.shape[0])

    def host_pattern(self):
        """(indptr, indices) as NumPy arrays owned by the caller."""
        cache = getattr(self.plan, 'host_pattern', None) if self.plan is not None else None
        if cache is None:
            cache = (to_host(self.indptr), to_host(self.indices))
            if self.plan is not None:
                self.plan.host_pattern = cache
        return cache[0].copy(), cache[1].copy()

    def to_scipy(self):
        from scipy.sparse import csr_matrix
        indptr, indices = self.host_pattern()
        return csr_matrix((to_host(self.data), indices, indptr), shape=self.shape)


class Plan:
    """Symbolic result for one (basis, form): the CSR pattern and the
    reduction plan (``perm``/``seg``), the bit mask of the exact-zero
    contributions it was built from (``mask``) and its serial number on the
    basis (``signature``)."""

    def __init__(self, indptr, indices, perm, seg, nnz, signature):
        self.indptr, self.indices, self.perm, self.seg = indptr, indices, perm, seg
        self.nnz, self.signature = nnz, signature
        self.fused = None   # fused-kernel plan, built on demand
        self.host_pattern = None   # (indptr, indices) on the host, after the first transfer
        self.mask = None    # exact-zero mask of the contributions
        self.nzeros = 0     # number of exact-zero contributions
        self.seen = 0       # exact-path calls that matched this pattern


# patterns (exact-zero sets) kept per basis
PLAN_CACHE = 8


class DeviceBasis:

    def __init__(self, basis):
        mesh, elem = basis.cell_mesh, basis.elem
        check_host_inputs(mesh.p, mesh.t, basis.element_dofs, basis.N)
        lib = _lib.load()
        self.lib = lib
        self.basis = basis
        self.kind = mesh.kind
        self.dim = mesh.dim
        self.E = mesh.nelements
        self.Q = basis.W.shape[0]
        self.nbfun = basis.Nbfun
        self.N = basis.N
        self.p = to_dev(mesh.p, np.float64)
        self.t = to_dev(mesh.t, np.int32)
        self.dofs = to_dev(basis.element_dofs, np.int32)
        phi, dphi, gdphi = basis.tables()
        self.W = to_dev(basis.W, np.float64)
        self.phi = to_dev(phi, np.float64)
        self.dphi = to_dev(dphi, np.float64)
        self.gdphi = None if gdphi is None else to_dev(gdphi, np.float64)
        self.geom = _lib.Geom(_KIND[self.kind], self.dim, mesh.nverts, self.Q,
                              self.E, mesh.p.shape[1], self.p.data_ptr(),
                              self.t.data_ptr(), self.W.data_ptr(),
                              0 if self.gdphi is None else self.gdphi.data_ptr())
        self.space = _lib.Space(elem.nbs, 1 if elem.vector else 0, self.nbfun, 0,
                                self.N, self.dofs.data_ptr(), self.phi.data_ptr(),
                                self.dphi.data_ptr())
        self._G = None
        self._metric = None
        self.p_version = 0
        self.plans = {}          # serial number -> Plan
        self.validated = {}      # (form, params, coefficient shape, p version) -> Plan
        self.validated_coef = {}  # same key -> the coefficient it was validated with
        self.fused_plans = {}    # plan serial -> FusedPlan
        self.fused_last = {}     # form -> the last plan its fused path used
        self.last_verdict = None
        self.unfusable = {}      # plan serial -> why the fused plan was refused
        self.row_plan = None
        self.rows_plans = {}     # plan serial -> RowsFusedPlan (P2 simplices)
        self.rows_last = {}      # form -> the last plan its rows path used
        self.rows_skip = {}      # form -> its last rows call failed the check
        self.fused_skip = {}     # form -> its last fused P1 call failed the check
        self.mixed_plans = {}    # id(test device basis) -> (it, {serial: Plan})

    # -- coordinates ---------------------------------------------------------
    def set_coordinates(self, p):
        """Replace the node coordinates (device tensor (dim, nnodes) float64);
        the mapping is recomputed on next use, plans are revalidated."""
        assert p.shape == self.p.shape and p.dtype == torch.float64
        self.p = p.contiguous()
        self.geom.p = self.p.data_ptr()
        self._G = None
        self._metric = None
        self.p_version += 1

    def rows_metric(self):
        """Element metrics of the fused P2 kernel (``femhip_rows_metric``),
        recomputed when the coordinates change."""
        if self._metric is None:
            G = self.geometry()
            m = torch.empty(max(8 * self.E, 2), dtype=torch.float64, device=self.p.device)
            _lib.check(self.lib.femhip_rows_metric(ctypes.byref(self.geom), _lib.ptr(G),
                                                   _lib.ptr(m), _lib.stream_ptr()))
            self._metric = m
        return self._metric

    def geometry(self):
        if self._G is None:
            n = self.lib.femhip_geometry_size(ctypes.byref(self.geom))
            if n < 0:
                _lib.check(_lib.EINVAL)
            G = torch.empty(max(n, 1), dtype=torch.float64, device=self.p.device)
            status = torch.zeros(1, dtype=torch.int32, device=self.p.device)
            _lib.check(self.lib.femhip_geometry(ctypes.byref(self.geom), _lib.ptr(G),
                                                _lib.ptr(status), _lib.stream_ptr()))
            if self.kind == 'hex' and int(status.item()) != 0:
                raise Exception("Zero Jacobian determinant")
            self._G = G
        return self._G

    # -- forms ---------------------------------------------------------------
    def form_struct(self, name, kappa=None, lam=0.0, mu=0.0):
        mode, kp = 0, None
        if name in ('kappa_laplace', 'coef_mass', 'coef_load'):
            kp = kappa
            mode = 1 if kp.dim() == 1 else 2
        fs = _lib.Form(_FORM[name], mode, 0 if kp is None else kp.data_ptr(),
                       float(lam), float(2. * mu))
        return fs, kp

    def local_bilinear(self, name, kappa=None, lam=0.0, mu=0.0, out=None):
        G = self.geometry()
        fs, keep = self.form_struct(name, kappa, lam, mu)
        if out is None:
            out = torch.empty(self.nbfun * self.nbfun * self.E, dtype=torch.float64,
                              device=self.p.device)
        _lib.check(self.lib.femhip_local_bilinear(
            ctypes.byref(self.geom), ctypes.byref(self.space), ctypes.byref(fs),
            _lib.ptr(G), _lib.ptr(out), _lib.stream_ptr()))
        return out

    def local_linear(self, name='unit_load', kappa=None):
        G = self.geometry()
        fs, _ = self.form_struct(name, kappa)
        out = torch.empty(self.nbfun * self.E, dtype=torch.float64, device=self.p.device)
        _lib.check(self.lib.femhip_local_linear(
            ctypes.byref(self.geom), ctypes.byref(self.space), ctypes.byref(fs),
            _lib.ptr(G), _lib.ptr(out), _lib.stream_ptr()))
        return out

    # -- symbolic ------------------------------------------------------------
    def _workspace(self, ncontrib):
        nb = ctypes.c_size_t(0)
        _lib.check(self.lib.femhip_pattern_workspace(ncontrib, ctypes.byref(nb)))
        return torch.empty(int(nb.value), dtype=torch.uint8, device=self.p.device), int(nb.value)

    def build_plan(self, contrib, serial=None, test=None):
        """Pattern and reduction plan of ``contrib``; ``test``: the device
        basis of a distinct test space (its DOFs index the rows)."""
        test = self if test is None else test
        C = self.nbfun * test.nbfun * self.E
        dev = self.p.device
        ws, wsb = self._workspace(C)
        indptr = torch.empty(test.N + 1, dtype=torch.int32, device=dev)
        indices = torch.empty(max(C, 1), dtype=torch.int32, device=dev)
        perm = torch.empty(max(C, 1), dtype=torch.int32, device=dev)
        seg = torch.empty(max(C, 1) + 1, dtype=torch.int32, device=dev)
        nnz = ctypes.c_int64(0)
        _lib.check(self.lib.femhip_pattern_build(
            _lib.ptr(test.dofs), _lib.ptr(self.dofs), test.nbfun, self.nbfun, self.E,
            test.N, self.N, _lib.ptr(contrib), _lib.ptr(indptr), _lib.ptr(indices),
            _lib.ptr(perm), _lib.ptr(seg), ctypes.byref(nnz), _lib.ptr(ws), wsb,
            _lib.stream_ptr()))
        n = int(nnz.value)
        del ws
        plan = Plan(indptr, indices[:n].clone(), perm, seg[:n + 1].clone(), n, serial)
        # contrib None: the connectivity pattern (no contribution pruned)
        if contrib is not None:
            plan.mask = self.zero_mask(contrib)
            # the kept contributions end at seg[nnz] (zeros sort last)
            plan.nzeros = C - int(plan.seg[n].item())
        return plan

    def zero_mask(self, contrib):
        """Exact-zero mask of the contributions (one bit each)."""
        n = contrib.numel()
        mask = torch.empty(max((n + 63) // 64, 1), dtype=torch.int64, device=self.p.device)
        _lib.check(self.lib.femhip_zero_mask(_lib.ptr(contrib), n, _lib.ptr(mask),
                                             _lib.stream_ptr()))
        return mask

    def zero_mismatch(self, contrib, mask):
        """Number of contributions whose exact-zero bit differs from mask."""
        out = torch.empty(1, dtype=torch.int64, device=self.p.device)
        _lib.check(self.lib.femhip_zero_mask_mismatch(_lib.ptr(contrib), contrib.numel(),
                                                      _lib.ptr(mask), _lib.ptr(out),
                                                      _lib.stream_ptr()))
        return int(out.item())

    def plan_for(self, contrib):
        """Plan whose exact-zero positions equal those of ``contrib``,
        compared bit for bit with each known plan's mask (most recently
        matched first; patterns are shared by every form with the same zero
        set on this basis), else a new one.  At most ``PLAN_CACHE`` patterns
        are kept per basis: a loop over inputs whose zero sets all differ
        (a coefficient with zeros, say) does not accumulate them."""
        serials = list(reversed(list(self.plans)))
        if serials:
            # every cached mask compared on the device, one read-back
            counts = torch.empty(len(serials), dtype=torch.int64, device=self.p.device)
            for k, serial in enumerate(serials):
                _lib.check(self.lib.femhip_zero_mask_mismatch(
                    _lib.ptr(contrib), contrib.numel(), _lib.ptr(self.plans[serial].mask),
                    _lib.ptr(counts[k:]), _lib.stream_ptr()))
            for k, c in enumerate(counts.cpu().tolist()):
                if c == 0:
                    serial = serials[k]
                    self.plans[serial] = self.plans.pop(serial)   # most recent last
                    return self.plans[serial]
        serial = getattr(self, '_serial', len(self.plans))
        self._serial = serial + 1
        plan = self.build_plan(contrib, serial)
        self.plans[serial] = plan
        if len(self.plans) > PLAN_CACHE:
            live = {p.signature for p in list(self.fused_last.values()) +
                    list(getattr(self, 'rows_last', {}).values())}
            for old in list(self.plans):
                if old != serial and old not in live:
                    del self.plans[old]
                    self.fused_plans.pop(old, None)
                    getattr(self, 'rows_plans', {}).pop(old, None)
                    self.unfusable.pop(old, None)
                    break
        return plan

    def fused_plan(self, plan, form='laplace'):
        """Fused plan for a CSR pattern (the configuration the autotuner
        cached for this mesh and pattern when a cache is named, else the
        default one)."""
        from .fused import FusedPlan
        from . import autotune
        fp = self.fused_plans.get(plan.signature)
        if fp is None:
            torch.cuda.current_stream().synchronize()
            ip, ix = plan.indptr.cpu().numpy(), plan.indices.cpu().numpy()
            cfg = None
            if autotune._cache_path(None):
                cfg = autotune.lookup(autotune.plan_key(self.kind, form, self.basis.cell_mesh.t,
                                                        ip, ix))
            if cfg is not None:
                try:
                    fp = FusedPlan(self.basis.cell_mesh, ip, ix, cfg[0], cfg[1], block=cfg[2],
                                   order=cfg[3])
                    fp.groups = cfg[4]
                except (NotImplementedError, ValueError):
                    fp = None    # a cached configuration this build refuses
            if fp is None:
                fp = FusedPlan(self.basis.cell_mesh, ip, ix)
            self.fused_plans[plan.signature] = fp
        return fp

    def rows_plan(self):
        if self.row_plan is None:
            C = self.nbfun * self.E
            dev = self.p.device
            ws, wsb = self._workspace(C)
            perm = torch.empty(max(C, 1), dtype=torch.int32, device=dev)
            seg = torch.empty(self.N + 1, dtype=torch.int32, device=dev)
            _lib.check(self.lib.femhip_rows_plan_build(
                _lib.ptr(self.dofs), self.nbfun, self.E, self.N, _lib.ptr(perm),
                _lib.ptr(seg), _lib.ptr(ws), wsb, _lib.stream_ptr()))
            torch.cuda.current_stream().synchronize()
            self.row_plan = (perm, seg)
        return self.row_plan

    def assemble_mixed(self, test, name, kappa=None, lam=0.0, mu=0.0):
        """CSR assembly with trial functions from this basis and test
        functions from ``test`` (another device basis on the same elements
        and quadrature points; ``bilinear_form.py:58-128``): exact-order
        local entries, a pattern per exact-zero set (matched bit for bit as
        in ``plan_for``), segmented sums; shape (test.N, self.N)."""
        G = self.geometry()
        fs, keep = self.form_struct(name, kappa, lam, mu)
        C = self.nbfun * test.nbfun * self.E
        dev = self.p.device
        if C == 0:
            return DeviceCSR(torch.zeros(test.N + 1, dtype=torch.int32, device=dev),
                             torch.zeros(0, dtype=torch.int32, device=dev),
                             torch.zeros(0, dtype=torch.float64, device=dev),
                             (test.N, self.N))
        contrib = torch.empty(C, dtype=torch.float64, device=dev)
        _lib.check(self.lib.femhip_local_bilinear_mixed(
            ctypes.byref(self.geom), ctypes.byref(self.space), ctypes.byref(test.space),
            ctypes.byref(fs), _lib.ptr(G), _lib.ptr(contrib), _lib.stream_ptr()))
        holder = self.mixed_plans.setdefault(id(test), (test, {}))
        plans = holder[1]
        plan = None
        for serial in sorted(plans, reverse=True):
            if self.zero_mismatch(contrib, plans[serial].mask) == 0:
                plan = plans[serial]
                break
        if plan is None:
            plan = self.build_plan(contrib, len(plans), test=test)
            plans[plan.signature] = plan
        data = self.segment_sum(contrib, plan.perm, plan.seg, plan.nnz)
        return DeviceCSR(plan.indptr, plan.indices, data, (test.N, self.N), plan)

    def segment_sum(self, contrib, perm, seg, nseg, out=None):
        if out is None:
            out = torch.empty(max(nseg, 1), dtype=torch.float64, device=self.p.device)[:nseg]
        _lib.check(self.lib.femhip_segment_sum(_lib.ptr(contrib), _lib.ptr(perm),
                                               _lib.ptr(seg), nseg, _lib.ptr(out),
                                               _lib.stream_ptr()))
        return out

    # -- assembly ------------------------------------------------------------
    def _validated_plan(self, vkey, kappa):
        """The plan validated for these inputs, or None: the key holds the
        form, its parameters and the coordinate version; a coefficient must
        equal (bit for bit, on the device) the one the plan was validated
        with."""
        plan = self.validated.get(vkey)
        if plan is None:
            return None
        ref = self.validated_coef.get(vkey)
        if kappa is None:
            return plan if ref is None else None
        if ref is None or not torch.equal(ref, kappa):
            return None
        return plan

    def _validate(self, vkey, plan, kappa):
        self.validated[vkey] = plan
        self.validated_coef[vkey] = None if kappa is None else kappa.detach().clone()

    def fused_try(self, name, kappa=None, load=None):
        """The fused kernel on the last pattern validated for ``name``, the
        call itself validated against that pattern: the CSR values
        (DeviceCSR) when this input prunes exactly the plan's entries,
        else None (the verdict is kept in ``last_verdict``)."""
        plan = self.fused_last.get(name)
        if plan is None:
            return None
        fp = self.fused_plans[plan.signature]
        out = torch.empty(max(plan.nnz, 1), dtype=torch.float64,
                          device=self.p.device)[:plan.nnz]
        chk = fp.check_struct(self, plan.mask, plan.nzeros)
        fp.assemble(name, self.p, out, kappa, load=load, check=chk, groups=fp.groups)
        v = fp.verdict.cpu().tolist()      # synchronises
        self.last_verdict = v
        if v[0] != 1:
            return None
        return DeviceCSR(plan.indptr, plan.indices, out, (self.N, self.N), plan)

    def rows_fused_try(self, name, kappa=None):
        """The fused rows kernel (P2 simplices) on the last pattern validated
        for ``name``, checked against it in the same pass: the CSR values
        (DeviceCSR) when this input's exact zeros are the pattern's, else
        None (verdict in ``last_verdict``)."""
        plan = self.rows_last.get(name)
        if plan is None:
            return None
        fp = self.rows_plans[plan.signature]
        out = torch.empty(max(plan.nnz, 1), dtype=torch.float64,
                          device=self.p.device)[:plan.nnz]
        fp.assemble(name, self, out, kappa)
        v = fp.verdict.cpu().tolist()      # synchronises
        if v[0] != 1 and (v[3] & _lib.FUSED_OVERFLOW):
            fp.grow_queue()                # more entries near zero than queued
            fp.assemble(name, self, out, kappa)
            v = fp.verdict.cpu().tolist()
        self.last_verdict = v
        if v[0] != 1:
            return None
        return DeviceCSR(plan.indptr, plan.indices, out, (self.N, self.N), plan)

    def rows_plan_for(self, plan, name):
        """Rows plan of a validated P2 pattern (built once per pattern);
        None when the mesh cannot be encoded (the exact path stays)."""
        from .fused import RowsFusedPlan
        if plan.signature in self.unfusable:
            return None
        fp = self.rows_plans.get(plan.signature)
        if fp is None:
            torch.cuda.current_stream().synchronize()
            try:
                fp = RowsFusedPlan(self.basis, plan.indptr.cpu().numpy(),
                                   plan.indices.cpu().numpy())
            except (NotImplementedError, ValueError, OverflowError) as exc:
                self.unfusable[plan.signature] = str(exc)
                return None
            fp.set_pattern_zeros(plan.mask, plan.nzeros)
            self.rows_plans[plan.signature] = fp
        self.rows_last[name] = plan
        return fp

    def assemble_bilinear(self, name, kappa=None, lam=0.0, mu=0.0, path='auto'):
        """CSR assembly of a built-in form.

        P1 triangles/tetrahedra (laplace, mass, per-element kappa_laplace):
        the fused kernel on the last pattern validated for the form, the
        call validated in the same launch sequence (entries within the
        rounding bound of zero recomputed in the reference's arithmetic and
        compared with the pattern's zero set); when the check fails, or on
        the first call, the exact path (local entries in the reference's
        arithmetic -> pattern matched by its exact-zero mask or built ->
        segmented sums).  A pattern gets its fused plan the first time the
        form has none, else when it is matched a second time; after a failed
        check the form's calls take the exact path until one matches a
        pattern that has a fused plan.  P2 simplices (laplace, mass,
        kappa_laplace per element) do the same with the fused rows kernel.

        Vector Q2-hex elasticity (MFMA kernel) and the other scalar forms
        (exact path without the zero-mask pass) reuse the pattern validated
        for the same form, parameters, coordinates and coefficient (compared
        bit for bit on the device).
        """
        from . import fused
        if self.E == 0:   # an empty element subset: the all-zero matrix
            self.local_bilinear(name, kappa, lam, mu)   # form check only
            dev = self.p.device
            return DeviceCSR(torch.zeros(self.N + 1, dtype=torch.int32, device=dev),
                             torch.zeros(0, dtype=torch.int32, device=dev),
                             torch.zeros(0, dtype=torch.float64, device=dev), (self.N, self.N))
        fusable = (path != 'generic' and fused.eligible(self.basis, name)
                   and (kappa is None or kappa.dim() == 1))
        mfma = (path != 'generic' and name == 'linear_elasticity' and self.kind == 'hex'
                and self.basis.elem.vector and self.basis.elem.nbs == 27 and self.Q == 27)
        # the other scalar forms: the exact path without the exact-zero
        # comparison (same inputs, same zeros)
        scalar = (path != 'generic' and not fusable and not self.basis.elem.vector
                  and name in ('laplace', 'mass', 'kappa_laplace', 'coef_mass'))
        rows = (path != 'generic' and (kappa is None or kappa.dim() == 1)
                and fused.rows_eligible(self.basis, name))
        vkey = (name, lam, mu, None if kappa is None else tuple(kappa.shape), self.p_version)
        if fusable and not self.fused_skip.get(name):
            A = self.fused_try(name, kappa)
            if A is not None:
                self.validated[vkey] = A.plan
                return A
            # the input's zeros differ from the pattern's: the next calls go
            # to the exact path until one matches a pattern with a fused plan
            self.fused_skip[name] = True
        if rows and not self.rows_skip.get(name):
            A = self.rows_fused_try(name, kappa)
            if A is not None:
                self.validated[vkey] = A.plan
                return A
            # the input's zeros differ from the pattern's: the next calls go
            # straight to the exact path until one matches a rows pattern
            self.rows_skip[name] = True
        plan = self._validated_plan(vkey, kappa) if (mfma or scalar) else None
        if plan is not None and rows and plan.signature not in self.unfusable:
            # the same input as a validated call (e.g. after the zeros
            # changed once): its pattern gets a rows plan when it comes back
            fp = self.rows_plans.get(plan.signature)
            if fp is None:
                plan.seen += 1
                if plan.seen >= 2:
                    fp = self.rows_plan_for(plan, name)
            if fp is not None:
                self.rows_last[name] = plan
                self.rows_skip[name] = False
                A = self.rows_fused_try(name, kappa)
                if A is not None:
                    return A
        if plan is not None:
            out = torch.empty(max(plan.nnz, 1), dtype=torch.float64,
                              device=self.p.device)[:plan.nnz]
            if mfma:
                self.elasticity_mfma(lam, mu, plan, out)
            else:
                self.segment_sum(self.local_bilinear(name, kappa), plan.perm, plan.seg,
                                 plan.nnz, out=out)
            return DeviceCSR(plan.indptr, plan.indices, out, (self.N, self.N), plan)
        contrib = self.local_bilinear(name, kappa, lam, mu)
        plan = self.plan_for(contrib)
        data = self.segment_sum(contrib, plan.perm, plan.seg, plan.nnz)
        if fusable and plan.signature not in self.unfusable:
            plan.seen += 1
            if plan.signature in self.fused_plans or name not in self.fused_last \
                    or plan.seen >= 2:
                try:
                    self.fused_plan(plan, name)
                    self.fused_last[name] = plan
                    self.fused_skip[name] = False
                    self.validated[vkey] = plan
                except (NotImplementedError, ValueError) as exc:
                    # a mesh the fused plan cannot encode (e.g. a CSR row
                    # longer than 31 entries): this pattern stays on the
                    # exact path, which is always available
                    self.unfusable[plan.signature] = str(exc)
        elif mfma or scalar:
            self._validate(vkey, plan, kappa)
            if rows and plan.signature not in self.unfusable:
                # a rows plan (host build) when the form has none yet, or
                # when this pattern comes back: inputs whose zero sets keep
                # changing stay on the exact path without a build per call
                plan.seen += 1
                if plan.signature in self.rows_plans or name not in self.rows_last \
                        or plan.seen >= 2:
                    if self.rows_plan_for(plan, name) is not None:
                        self.rows_skip[name] = False
        return DeviceCSR(plan.indptr, plan.indices, data, (self.N, self.N), plan)

    def elasticity_mfma(self, lam, mu, plan, out, loc=None):
        """Numeric vector-Q2-hex elasticity: MFMA local matrices (element
        major, row-major) + row sums over the plan (each local row read once,
        contiguous; ``femhip_row_sum_em``), or the slot-wise segmented sums
        when a row is longer than the row kernel takes."""
        G = self.geometry()
        nb = self.nbfun
        if loc is None:
            loc = torch.empty(nb * nb * self.E, dtype=torch.float64, device=self.p.device)
        _lib.check(self.lib.femhip_hex2_elasticity_mfma(
            ctypes.byref(self.geom), ctypes.byref(self.space), float(lam), float(mu),
            _lib.ptr(G), _lib.ptr(loc), 0, _lib.stream_ptr()))
        rs = self._row_sum_plan(plan)
        if rs is None:
            _lib.check(self.lib.femhip_segment_sum_em(
                _lib.ptr(loc), _lib.ptr(plan.perm), _lib.ptr(plan.seg), plan.nnz, self.E, nb,
                _lib.ptr(out), _lib.stream_ptr()))
        else:
            pos, rowe_ptr, rowe, max_len = rs
            _lib.check(self.lib.femhip_row_sum_em(
                _lib.ptr(loc), _lib.ptr(pos), _lib.ptr(rowe_ptr), _lib.ptr(rowe),
                _lib.ptr(plan.indptr), self.N, nb, max_len, _lib.ptr(out), _lib.stream_ptr()))
        return out

    def _row_sum_plan(self, plan):
        """(pos, rowe_ptr, rowe, max row length) of a pattern for the row
        sums, built once per plan; None for rows over 384 slots."""
        rs = getattr(plan, 'row_sums', 0)
        if rs != 0:
            return rs
        dev = self.p.device
        lens = plan.indptr[1:] - plan.indptr[:-1]
        max_len = int(lens.max().item()) if self.N > 0 else 0
        if max_len > 384:
            plan.row_sums = None
            return None
        nb = self.nbfun
        # local rows lr = e*nb + i grouped by their global row, increasing e
        rows_of_lr = self.dofs.t().contiguous().reshape(-1).long()   # [E][nb] -> lr
        order = torch.argsort(rows_of_lr, stable=True)
        counts = torch.bincount(rows_of_lr, minlength=self.N)
        rowe_ptr = torch.zeros(self.N + 1, dtype=torch.int32, device=dev)
        rowe_ptr[1:] = torch.cumsum(counts, 0).to(torch.int32)
        rowe = order.to(torch.int32).contiguous()
        pos = torch.full((self.E * nb * nb,), -1, dtype=torch.int16, device=dev)
        _lib.check(self.lib.femhip_row_sum_em_plan(
            _lib.ptr(plan.perm), _lib.ptr(plan.seg), _lib.ptr(plan.indptr), self.N, self.E, nb,
            _lib.ptr(pos), _lib.stream_ptr()))
        plan.row_sums = (pos, rowe_ptr, rowe, max_len)
        return plan.row_sums

    def assemble_linear(self, name='unit_load', kappa=None):
        if self.E == 0:   # an empty element subset: the zero vector
            self.local_linear(name, kappa)
            return torch.zeros(self.N, dtype=torch.float64, device=self.p.device)
        contrib = self.local_linear(name, kappa)
        perm, seg = self.rows_plan()
        return self.segment_sum(contrib, perm, seg, self.N)

    # -- default form parameters -------------------------------------------
    def global_coordinates(self):
        """(dim, E, Q) coordinates of the quadrature points."""
        X = to_dev(self.basis.X, np.float64)
        phi1 = None
        if self.kind == 'hex':
            from .element import ElementHex1
            phi1 = to_dev(ElementHex1().tables(self.basis.X)[0], np.float64)
        out = torch.empty((self.dim, self.E, self.Q), dtype=torch.float64, device=self.p.device)
        _lib.check(self.lib.femhip_global_coordinates(
            ctypes.byref(self.geom), _lib.ptr(X), _lib.ptr(phi1), _lib.ptr(out),
            _lib.stream_ptr()))
        torch.cuda.current_stream().synchronize()
        return out

    def mesh_parameters(self):
        """(E, Q) mesh parameter |detDF| ** (1 / dim)."""
        G = self.geometry()
        out = torch.empty((self.E, self.Q), dtype=torch.float64, device=self.p.device)
        _lib.check(self.lib.femhip_mesh_parameters(ctypes.byref(self.geom), _lib.ptr(G),
                                                   _lib.ptr(out), _lib.stream_ptr()))
        return out


class FacetDeviceBasis(DeviceBasis):
    """Device side of a ``FacetBasis``: the cell mesh (for the element
    mapping of the owning elements), the facets, ``tind`` and the facet
    quadrature.  Local entries come from ``femhip_facet_local``; the pattern
    and the sums go through the generic path of ``DeviceBasis`` with one
    "element" per facet (``element_dofs[:, tind]``)."""

    def __init__(self, basis):
        mesh, elem = basis.mesh, basis.elem
        check_host_inputs(mesh.p, mesh.t, basis.element_dofs, basis.N)
        lib = _lib.load()
        self.lib = lib
        self.basis = basis
        self.kind = mesh.kind
        self.dim = mesh.dim
        self.E = basis.nelems              # one local block per facet
        self.Q = basis.W.shape[0]
        self.nbfun = basis.Nbfun
        self.N = basis.N
        self.p = to_dev(mesh.p, np.float64)
        self.t = to_dev(mesh.t, np.int32)
        self.dofs = to_dev(basis.element_dofs, np.int32)
        self.facets = to_dev(mesh.facets[:, basis.find], np.int32)
        self.tind = to_dev(basis.tind, np.int32)
        self.X = to_dev(basis.X, np.float64)
        self.W = to_dev(basis.W, np.float64)
        self.fgeom = _lib.FacetGeom(_KIND[self.kind], self.dim, elem.maxdeg, self.Q,
                                    self.E, mesh.p.shape[1], mesh.nelements,
                                    self.p.data_ptr(), self.facets.data_ptr(),
                                    self.t.data_ptr(), self.tind.data_ptr(),
                                    self.X.data_ptr(), self.W.data_ptr())
        self._G = None
        self._normals = None
        self.p_version = 0
        self.plans = {}
        self.validated = {}
        self.validated_coef = {}
        self.fused_plans = {}
        self.fused_last = {}
        self.last_verdict = None
        self.unfusable = {}
        self.row_plan = None

    def set_coordinates(self, p):
        assert p.shape == self.p.shape and p.dtype == torch.float64
        self.p = p.contiguous()
        self.fgeom.p = self.p.data_ptr()
        self._normals = None
        self.p_version += 1

    def geometry(self):
        raise NotImplementedError("a facet basis maps its facets inside the facet kernels")

    def _local(self, name, kappa, out):
        if self.E == 0:
            return out
        mode = 0 if kappa is None else (1 if kappa.dim() == 1 else 2)
        _lib.check(self.lib.femhip_facet_local(ctypes.byref(self.fgeom), _FORM[name],
                                               _lib.ptr(kappa), mode, _lib.ptr(out),
                                               _lib.stream_ptr()))
        return out

    def local_bilinear(self, name, kappa=None, lam=0.0, mu=0.0, out=None):
        if name not in ('mass', 'coef_mass'):
            raise NotImplementedError("facet bilinear forms: mass and coef_mass")
        if out is None:
            out = torch.empty(self.nbfun * self.nbfun * self.E, dtype=torch.float64,
                              device=self.p.device)
        return self._local(name, kappa if name == 'coef_mass' else None, out)

    def local_linear(self, name='unit_load', kappa=None):
        if name not in ('unit_load', 'coef_load'):
            raise NotImplementedError("facet linear forms: unit_load and coef_load")
        out = torch.empty(self.nbfun * self.E, dtype=torch.float64, device=self.p.device)
        return self._local(name, kappa if name == 'coef_load' else None, out)

    def interpolate(self, w, grad=True):
        """(value (nf, Q), grad (dim, nf, Q) or None) of a DOF vector
        (device tensor) at the facet quadrature points."""
        value = torch.empty((self.E, self.Q), dtype=torch.float64, device=self.p.device)
        gr = (torch.empty((self.dim, self.E, self.Q), dtype=torch.float64, device=self.p.device)
              if grad else None)
        if self.E:
            _lib.check(self.lib.femhip_facet_interpolate(
                ctypes.byref(self.fgeom), _lib.ptr(self.dofs), _lib.ptr(w), _lib.ptr(value),
                _lib.ptr(gr), _lib.stream_ptr()))
        return value, gr

    def global_coordinates(self):
        """(x (dim, nf, Q), h (nf, Q)): facet quadrature points and the mesh
        parameter."""
        x = torch.empty((self.dim, self.E, self.Q), dtype=torch.float64, device=self.p.device)
        h = torch.empty((self.E, self.Q), dtype=torch.float64, device=self.p.device)
        if self.E:
            _lib.check(self.lib.femhip_facet_global_coordinates(
                ctypes.byref(self.fgeom), _lib.ptr(x), _lib.ptr(h), _lib.stream_ptr()))
        return x, h

    def normals(self):
        """(dim, nfacets, Q) outward unit normals seen from ``f2t[0]``."""
        if self._normals is None and self.E == 0:
            self._normals = torch.zeros((self.dim, 0, self.Q), dtype=torch.float64,
                                        device=self.p.device)
        if self._normals is None:
            tn = to_dev(self.basis.tind_normals, np.int32)
            floc = to_dev(self.basis.facet_local_index(), np.int32)
            out = torch.empty((self.dim, self.E, self.Q), dtype=torch.float64,
                              device=self.p.device)
            _lib.check(self.lib.femhip_facet_normals(
                ctypes.byref(self.fgeom), _lib.ptr(tn), _lib.ptr(floc), _lib.ptr(out),
                _lib.stream_ptr()))
            torch.cuda.current_stream().synchronize()
            self._normals = out
        return self._normals

    def assemble_bilinear(self, name, kappa=None, lam=0.0, mu=0.0, path='auto'):
        """Exact path: local entries -> pattern (matched by exact-zero mask)
        -> segmented sums."""
        if self.E == 0:
            self.local_bilinear(name, kappa)   # form check only
            dev = self.p.device
            return DeviceCSR(torch.zeros(self.N + 1, dtype=torch.int32, device=dev),
                             torch.zeros(0, dtype=torch.int32, device=dev),
                             torch.zeros(0, dtype=torch.float64, device=dev), (self.N, self.N))
        contrib = self.local_bilinear(name, kappa)
        plan = self.plan_for(contrib)
        data = self.segment_sum(contrib, plan.perm, plan.seg, plan.nnz)
        return DeviceCSR(plan.indptr, plan.indices, data, (self.N, self.N), plan)

    def assemble_linear(self, name='unit_load', kappa=None):
        if self.E == 0:
            self.local_linear(name, kappa)
            return torch.zeros(self.N, dtype=torch.float64, device=self.p.device)
        return super().assemble_linear(name, kappa)
