"""Autotuner of the fused P1 kernel's launch configuration.

Tunables of ``FusedPlan`` / ``femhip_fused_assemble``:

* the LDS budget per workgroup (CSR slots and local nodes) -- the partition
  size, i.e. the elements each workgroup assembles, and how many
  workgroups fit on a CU;
* the workgroup size (elements per chunk: 512 or 1024);
* the record order inside a partition (colour classes or element order);
* the grid: one workgroup per partition, or a persistent grid of ``groups``
  workgroups walking the partitions with a fixed stride.

The quadrature loops need no run-time choice: P1 gradients are constant per
element (the fused kernel has no point loop) and the exact kernels unroll
their point loop at compile time for each (element, number of points).

``tune_fused`` times every candidate on the caller's basis and pattern and
installs the fastest plan on the basis (later fused calls use it, with its
grid).  Results are written to and read from a JSON cache only when one is
named (``cache=`` or ``FEMHIP_AUTOTUNE_CACHE``); entries are keyed by the
element kind, the form and a SHA-256 of the mesh connectivity and the CSR
pattern, so a cached configuration is only ever applied to the mesh and
pattern it was measured on (the partitioning changes the fixed-point
rounding of the sums, within the stated tolerance, so results depend on the
plan the cache selects).
"""
import hashlib
import json
import os

import numpy as np
import torch

from . import fused as _fused

CANDIDATES = [
    # (slot_cap, node_cap, block, order, groups)
    (4096, 768, 512, 1, 0),
    (4096, 768, 512, 0, 0),
    (4096, 768, 512, 1, 768),
    (6144, 1152, 512, 1, 0),
    (6144, 1152, 512, 1, 512),
    (3072, 576, 512, 1, 0),
    (8192, 1536, 1024, 1, 0),
]


def _cache_path(cache):
    return cache or os.environ.get('FEMHIP_AUTOTUNE_CACHE')


def plan_key(kind, form, t, indptr, indices):
    h = hashlib.sha256()
    for a in (t, indptr, indices):
        h.update(np.ascontiguousarray(a).tobytes())
    return '%s/%s/%s' % (kind, form, h.hexdigest())


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(key, cache=None):
    """Cached configuration (slot_cap, node_cap, block, order, groups) of a
    mesh/pattern key, or None (also when no cache is named)."""
    path = _cache_path(cache)
    if not path:
        return None
    v = _load(path).get(key)
    return None if v is None else tuple(v['best'])


def _store(path, key, rec):
    db = _load(path)
    db[key] = rec
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, 'w') as f:
            json.dump(db, f, indent=1)
    except OSError:
        pass


def _time(fn, steps):
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / steps


def tune_fused(basis, form='laplace', candidates=None, steps=20, cache=None):
    """Time the fused kernel for every candidate configuration on ``basis``
    (its validated pattern for ``form``), install the fastest plan on the
    basis and, when a cache is named, store it there.  Returns {'best':
    config, 'ms': ..., 'trials': [[config, ms, dup], ...]}."""
    from .models import laplace, mass
    f = {'laplace': laplace, 'mass': mass}[form]
    d = basis.device()
    f.assemble_device(basis)
    plan = d.fused_last.get(form)
    if plan is None:
        raise NotImplementedError("no fused pattern for %s on this basis" % form)
    ip, ix = plan.indptr.cpu().numpy(), plan.indices.cpu().numpy()
    out = torch.empty(plan.nnz, dtype=torch.float64, device=d.p.device)
    ref = None
    trials, best = [], None
    for cfg in (candidates or CANDIDATES):
        sc, nc, block, order, groups = cfg
        try:
            fp = _fused.FusedPlan(basis.cell_mesh, ip, ix, sc, nc, block=block, order=order)
        except (NotImplementedError, ValueError):   # caps the mesh cannot satisfy
            continue
        if fp.lds_bytes() > 160 * 1024:
            continue
        ms = _time(lambda: fp.assemble(form, d.p, out, groups=groups), steps)
        if ref is None:
            ref = out.clone()
        else:   # every configuration computes the same matrix (to rounding)
            err = float((out - ref).abs().max() / ref.abs().max())
            assert err <= 1e-12, (cfg, err)
        trials.append((list(cfg), ms, fp.info.dup))
        if best is None or ms < best[1]:
            best = (cfg, ms, fp)
    if best is None:
        raise RuntimeError("no candidate configuration fits")
    fp = best[2]
    fp.groups = best[0][4]
    d.fused_plans[plan.signature] = fp
    rec = {'best': list(best[0]), 'ms': best[1], 'trials': trials}
    path = _cache_path(cache)
    if path:
        _store(path, plan_key(d.kind, form, basis.cell_mesh.t, ip, ix), rec)
    return rec
