"""Autotuner of the fused P1 kernel's launch parameters.

Tunables: the LDS budget per workgroup (CSR slots and local nodes), which
sets the partition size -- elements per workgroup -- and the number of
workgroups resident per CU; and the number of persistent workgroups.  The
quadrature loops of the exact kernels are unrolled at compile time per
(element, number of points), so they need no run-time choice.

Results are cached in JSON (``FEMHIP_AUTOTUNE_CACHE``, default
``~/.cache/femhip/autotune.json``) keyed by element kind, form and mesh-size
bucket; ``DeviceBasis.fused_plan`` consults the cache.
"""
import json
import math
import os

import torch

from . import fused as _fused

CACHE = os.environ.get('FEMHIP_AUTOTUNE_CACHE',
                       os.path.join(os.path.expanduser('~'), '.cache', 'femhip',
                                    'autotune.json'))

CANDIDATES = [(4096, 768), (6144, 1024), (8192, 1280), (9000, 1500), (12000, 2000)]


def _key(kind, form, nnodes):
    return '%s/%s/%d' % (kind, form, int(math.log2(max(nnodes, 1))))


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(kind, form, nnodes):
    """Cached (slot_cap, node_cap, groups) or None."""
    v = _load().get(_key(kind, form, nnodes))
    return None if v is None else tuple(v['best'])


def _store(key, rec):
    db = _load()
    db[key] = rec
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, 'w') as f:
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


def tune_fused(basis, form='laplace', candidates=None, groups=(0,), steps=20):
    """Time the fused kernel for every candidate LDS budget (and persistent
    group count) on ``basis``; keep and cache the fastest.  Returns
    {'best': (slot_cap, node_cap, groups), 'ms': ..., 'trials': [...]}."""
    from .models import laplace, mass
    f = {'laplace': laplace, 'mass': mass}[form]
    d = basis.device()
    f.assemble_device(basis)
    plan = next(p for k, p in d.validated.items() if k[0] == form)
    ip, ix = plan.indptr.cpu().numpy(), plan.indices.cpu().numpy()
    out = torch.empty(plan.nnz, dtype=torch.float64, device=d.p.device)
    trials, best = [], None
    for sc, nc in (candidates or CANDIDATES):
        try:
            fp = _fused.FusedPlan(basis.mesh, ip, ix, sc, nc)
        except Exception:   # caps the mesh cannot satisfy
            continue
        for g in groups:
            ms = _time(lambda: fp.assemble(form, d.p, out, groups=g), steps)
            trials.append(((sc, nc, g), ms, fp.info.dup))
            if best is None or ms < best[1]:
                best = ((sc, nc, g), ms, fp)
    if best is None:
        raise RuntimeError("no candidate configuration fits")
    d.fused_plans[plan.signature] = best[2]
    rec = {'best': list(best[0]), 'ms': best[1],
           'trials': [[list(c), ms, dup] for c, ms, dup in trials]}
    _store(_key(basis.mesh.kind, form, basis.N), rec)
    return rec
