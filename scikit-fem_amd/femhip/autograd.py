"""Differentiable P1 assembly: ``torch.autograd.Function`` whose backward is
the hand-written fused adjoint kernel (``femhip_fused_adjoint``).

    data, b, (indptr, indices) = kappa_laplace_with_load(basis, kappa, p)

``data`` holds the kappa-Laplace CSR values on the *connectivity* pattern
of the mesh (every (row, col) pair coupled by an element, no exact-zero
pruning: a slot that happens to be 0 for these coordinates still has a
nonzero derivative with respect to them), ``b`` the unit-load vector.  The
backward maps (dL/ddata, dL/db) to dL/dkappa (per element) and dL/dp (node
coordinates) in one pass over the forward's fused plan, recomputing the
element geometry instead of storing it; the node sums are formed in the
plan's partitions (deterministic fixed point, no per-element array).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import fused as _fused


class _Ctx:
    """Per-basis state of the differentiable path: the connectivity pattern,
    its fused plan and the CSR slot of each row's diagonal."""

    def __init__(self, basis):
        d = basis.device()
        if not (_fused.eligible(basis, 'kappa_laplace')):
            raise NotImplementedError("the adjoint supports P1 triangles/tetrahedra")
        self.d = d
        self.plan = d.build_plan(None)            # connectivity pattern
        indptr = self.plan.indptr.cpu().numpy()
        indices = self.plan.indices.cpu().numpy()
        self.fp = _fused.FusedPlan(basis.cell_mesh, indptr, indices)
        if self.fp.info.pruned:
            raise RuntimeError("the connectivity pattern lacks an element pair")
        n = len(indptr) - 1
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
        k = np.flatnonzero(indices == rows)
        if len(k) != n:
            raise RuntimeError("a row of the connectivity pattern has no diagonal")
        self.diag_slot = torch.tensor(k.astype(np.int32), device=d.p.device)


def _ctx(basis):
    if getattr(basis, '_adjoint_ctx', None) is None:
        basis._adjoint_ctx = _Ctx(basis)
    return basis._adjoint_ctx


class _KappaLaplaceLoad(torch.autograd.Function):

    @staticmethod
    def forward(ctx, kappa, p, basis):
        c = _ctx(basis)
        d = c.d
        kappa = kappa.detach().contiguous()
        p = p.detach().contiguous()
        out = torch.empty(c.plan.nnz, dtype=torch.float64, device=p.device)
        b = torch.empty(p.shape[1], dtype=torch.float64, device=p.device)
        # CSR values and the unit-load vector in one fused pass
        c.fp.assemble('kappa_laplace', p, out, kappa, load=b)
        ctx.save_for_backward(kappa, p)
        ctx.basis = basis
        return out, b

    @staticmethod
    def backward(ctx, gdata, gload):
        kappa, p = ctx.saved_tensors
        c = _ctx(ctx.basis)
        d = c.d
        gdata = None if gdata is None else gdata.contiguous()
        gload = None if gload is None else gload.contiguous()
        dk = (torch.empty(d.E, dtype=torch.float64, device=p.device)
              if ctx.needs_input_grad[0] else None)
        dX = torch.empty_like(p)
        _lib.check(d.lib.femhip_fused_adjoint(
            ctypes.byref(c.fp.dev), _lib.ptr(c.fp.state('adjoint')), d.dim, _lib.ptr(p),
            p.shape[1], _lib.ptr(c.plan.indptr), _lib.ptr(c.diag_slot), _lib.ptr(kappa),
            _lib.ptr(gdata), _lib.ptr(gload), _lib.ptr(dk), _lib.ptr(dX), _lib.stream_ptr()))
        return dk, dX if ctx.needs_input_grad[1] else None, None


def kappa_laplace_with_load(basis, kappa, p=None):
    """Differentiable (CSR values, load vector) of kappa-Laplace + unit load
    on a P1 basis.  ``kappa``: (E,) float64 cuda tensor; ``p``: (dim, Nv)
    float64 cuda tensor (defaults to the mesh coordinates)."""
    c = _ctx(basis)
    if p is None:
        p = c.d.p
    data, b = _KappaLaplaceLoad.apply(kappa, p, basis)
    return data, b, (c.plan.indptr, c.plan.indices)


def quadratic_loss(data, indptr, indices, u, f):
    """L = 1/2 u^T A u - f^T u for A given by CSR arrays (torch ops)."""
    n = indptr.numel() - 1
    counts = (indptr[1:] - indptr[:-1]).long()
    rows = torch.repeat_interleave(torch.arange(n, device=u.device), counts)
    return 0.5 * (data * u[rows] * u[indices.long()]).sum() - (f * u).sum()
