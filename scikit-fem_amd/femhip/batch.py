"""Batched assembly.

``KappaSamples`` -- many per-element coefficient samples kappa^(s) on one
mesh, assembled into one ``[S, nnz]`` value array that shares a single CSR
pattern: the connectivity pattern of the mesh (no value-dependent pruning,
so every sample's reference pattern is a subset of it; slots the reference
drops for a particular sample hold values that are 0 up to rounding).
``export`` turns one sample into the reference's pruned CSR matrix.

``MeshBatch`` -- many independent P1 meshes of uneven size assembled
(Laplace + unit load) by ONE launch of the fused kernel over the
concatenated plans, with persistent workgroups; the launch can be captured
in a HIP graph (``capture``) and replayed.
"""
import numpy as np
import torch

from . import _lib
from . import fused as _fused


class KappaSamples:
    """kappa-Laplace for many per-element coefficient samples on one basis."""

    def __init__(self, basis):
        d = basis.device()
        if basis.elem.vector:
            raise NotImplementedError("scalar elements only")
        self.d = d
        self.plan = d.build_plan(None)          # connectivity pattern
        self.K1 = d.local_bilinear('laplace')   # unit-coefficient local entries
        self.nnz = self.plan.nnz
        # the local entries in plan order and their elements (contiguous runs
        # per slot for the batched kernel)
        perm = self.plan.perm[:self.K1.numel()].long()
        self.K1p = self.K1.index_select(0, perm).contiguous()
        self.elem = torch.remainder(perm, d.E).to(torch.int32).contiguous()
        self.indptr, self.indices = self.plan.indptr, self.plan.indices

    def export(self, s, kappa_s, values):
        """Sample ``s`` as the reference's CSR matrix (``DeviceCSR``): the
        pattern is the one ``kappa_laplace.assemble`` returns for
        ``kappa_s`` (its exact-zero entries found in the reference's
        arithmetic), the values are row ``s`` of the batch output
        ``values`` gathered into it."""
        from .device import DeviceCSR
        d = self.d
        kappa_s = kappa_s.reshape(-1).to(torch.float64).contiguous()
        plan = d.plan_for(d.local_bilinear('kappa_laplace', kappa_s))
        keys_c = self._keys(self.indptr, self.indices)
        keys_p = self._keys(plan.indptr, plan.indices)
        pos = torch.searchsorted(keys_c, keys_p)
        data = values[s].index_select(0, pos)
        return DeviceCSR(plan.indptr, plan.indices, data, (d.N, d.N), plan)

    def _keys(self, indptr, indices):
        n = indptr.numel() - 1
        rows = torch.repeat_interleave(torch.arange(n, device=indptr.device),
                                       (indptr[1:] - indptr[:-1]).long())
        return rows * n + indices.long()

    def assemble(self, kappa, out=None):
        """kappa: (S, E) float64 cuda tensor -> values (S, nnz)."""
        kappa = kappa.contiguous()
        S, E = kappa.shape
        assert E == self.d.E and kappa.dtype == torch.float64
        if out is None:
            out = torch.empty((S, self.nnz), dtype=torch.float64, device=kappa.device)
        _lib.check(self.d.lib.femhip_batched_kappa_values(
            _lib.ptr(self.K1p), _lib.ptr(self.elem), _lib.ptr(self.plan.seg),
            self.nnz, E, _lib.ptr(kappa), S, _lib.ptr(out), _lib.stream_ptr()))
        return out

    def capture(self, kappa, out):
        """Capture ``assemble(kappa, out)`` in a HIP graph.  ``kappa`` (S, E)
        and ``out`` (S, nnz) become the graph's buffers: copy new samples
        into ``kappa`` and call ``replay`` (the launch reads and writes
        those addresses on every replay)."""
        assert kappa.is_contiguous() and out.is_contiguous()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.assemble(kappa, out)            # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.assemble(kappa, out)
        self.graph = g
        self.graph_buffers = (kappa, out)
        return g

    def replay(self):
        self.graph.replay()
        return self.graph_buffers[1]


class MeshBatch:
    """Laplace + unit load on many P1 meshes in one launch."""

    def __init__(self, bases, groups=None, caps=None):
        from .models import laplace
        plans, ps = [], []
        self.bases = list(bases)
        for b in self.bases:
            if not _fused.eligible(b, 'laplace'):
                raise NotImplementedError("MeshBatch needs P1 triangles or tetrahedra")
            dev = b.device()
            laplace.assemble_device(b)          # exact path: validates the pattern
            plan = next(p for k, p in dev.validated.items() if k[0] == 'laplace')
            # the block-partition variant: plans of any number of meshes
            # concatenate into one launch
            fp = _fused.FusedPlan(b.cell_mesh, plan.indptr.cpu().numpy(),
                                  plan.indices.cpu().numpy(), *(caps or ()))
            plans.append((plan, fp))
            ps.append(dev.p)
        self.patterns = [pl for pl, _ in plans]
        self.plan = _fused.FusedPlan.concat([fp for _, fp in plans])
        self.p = torch.cat(ps, dim=1).contiguous()
        self.data = torch.empty(self.plan.nnz, dtype=torch.float64, device=self.p.device)
        self.load = torch.empty(self.plan.nnodes, dtype=torch.float64, device=self.p.device)
        self.groups = 0 if groups is None else int(groups)
        self.graph = None

    def assemble(self):
        """One launch: all meshes' CSR values and load vectors."""
        self.plan.assemble('laplace', self.p, self.data, load=self.load, groups=self.groups)
        return self.data, self.load

    def capture(self):
        """Capture ``assemble`` in a HIP graph (replay with ``replay``)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.assemble()                      # warm-up outside the graph
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.assemble()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
        return self.data, self.load

    def split(self):
        """Per-mesh (indptr, indices, data, load) views of the outputs."""
        out = []
        for (off_node, off_nnz, _), pat, b in zip(self.plan.offsets, self.patterns, self.bases):
            n, nnz = b.N, pat.nnz
            out.append((pat.indptr, pat.indices, self.data[off_nnz:off_nnz + nnz],
                        self.load[off_node:off_node + n]))
        return out
