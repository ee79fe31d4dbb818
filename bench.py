"""Benchmark of the scikit-fem assembly hot path on one MI355X.

Workload (one "step"): FP64 P1 Laplace assembly on the tetrahedral mesh
``MeshTet.init_tensor(linspace(0, 1, 101)^3)`` -- 6,000,000 tetrahedra,
1,030,301 DoF -- with the interior nodes jittered by U(-0.2h, 0.2h)
(``default_rng(0)``, h = 1/100).  Inputs (coordinates, connectivity, the
symbolic plan) are resident in HBM before the timed region; a step is the
numeric assembly into the CSR value array with the reference's pattern.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line (see the keys below).  ``--gpus`` above 1 is rejected:
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'scikit-fem_amd'))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
METRIC = "elements/s & %HBM roofline, FP64 P1-tet 1M DoF; batched+adjoint vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_mesh(n, jitter):
    import femhip
    x = np.linspace(0, 1, n)
    m = femhip.MeshTet.init_tensor(x, x, x)
    if jitter:
        m = m.jittered(jitter, 1.0 / (n - 1))
    return m


def time_stream(fn, steps, stream):
    """Average ms of ``fn`` over ``steps`` launches, HIP events on ``stream``."""
    import torch
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / steps


def cpu_baseline(n_sample=81):
    """The oracle (NumPy restatement of the reference algorithm) on a
    bounded sample: Basis + laplace.assemble on the 81^3 jittered mesh."""
    from oracle import skfem_np as O
    m = build_mesh(n_sample, 0.2)
    om = O.MeshTet(m.p, m.t)
    t0 = time.perf_counter()
    ob = O.CellBasis(om, O.ElementTetP1())
    t1 = time.perf_counter()
    O.bilinear_assemble(O.form_laplace, ob)
    t2 = time.perf_counter()
    E = m.nelements
    return {"value": E / (t2 - t0), "unit": "elements/s", "cores": 1,
            "kind": "port",
            "sample": "oracle Basis+laplace on MeshTet 81^3 (%d tets), jittered; "
                      "basis %.2fs, assembly %.2fs; single-threaded NumPy/SciPy"
                      % (E, t1 - t0, t2 - t1),
            "host_cpus": os.cpu_count(),
            "host_affinity": len(os.sched_getaffinity(0))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--n', type=int, default=101)
    ap.add_argument('--path', default='auto', choices=['auto', 'generic', 'fused'])
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import femhip
    from femhip.models import laplace
    dev = torch.device('cuda')
    torch.cuda.init()

    log('building mesh n=%d' % args.n)
    m = build_mesh(args.n, 0.2)
    basis = femhip.Basis(m, femhip.ElementTetP1())
    E, Nv, N = m.nelements, m.p.shape[1], basis.N

    # first call: upload + mapping + exact local entries + symbolic phase
    # (+ the fused plan on the host)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A = laplace.assemble_device(basis, path=args.path)
    torch.cuda.synchronize()
    first_call_ms = (time.perf_counter() - t0) * 1e3
    nnz = A.nnz
    log('first call %.1f ms, nnz %d' % (first_call_ms, nnz))

    dvb = basis.device()
    stream = torch.cuda.current_stream()
    path = 'fused' if (args.path != 'generic' and dvb.validated) else 'generic'
    extra = {}
    if path == 'fused':
        out = torch.empty(nnz, dtype=torch.float64, device=dev)
        plan = next(iter(dvb.validated.values()))
        fp = dvb.fused_plans[plan.signature]
        extra['fused_plan'] = fp.stats()

        def step():
            fp.assemble('laplace', dvb.p, out)
        kernel_name = 'p1_fused_kernel'
    else:
        plan = next(iter(dvb.plans.values()))
        contrib = torch.empty(16 * E, dtype=torch.float64, device=dev)
        out = torch.empty(nnz, dtype=torch.float64, device=dev)

        def step():
            dvb.local_bilinear('laplace', out=contrib)
            dvb.segment_sum(contrib, plan.perm, plan.seg, nnz, out=out)
        kernel_name = 'local_bilinear_kernel + segment_sum_kernel'

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ms = time_stream(step, args.steps, stream)
    torch.cuda.synchronize()

    # correctness of the timed output against the first-call result
    ref = A.data
    err = float((out - ref).abs().max() / ref.abs().max())

    # end-to-end (numeric + D2H + csr_matrix)
    t0 = time.perf_counter()
    step()
    from scipy.sparse import csr_matrix
    torch.cuda.synchronize()
    _ = csr_matrix((out.cpu().numpy(), A.indices.cpu().numpy(), A.indptr.cpu().numpy()),
                   shape=(N, N))
    e2e_ms = (time.perf_counter() - t0) * 1e3

    alg_bytes = 16 * E + 24 * Nv + 8 * nnz
    achieved = alg_bytes / (ms * 1e-3)
    res = {
        "metric": METRIC,
        "value": E / (ms * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (MeshTet.init_tensor 101^3, interior jitter U(-0.2h,0.2h) seed 0)",
        "config": {"workload": "P1-tet Laplace assembly, %d tets, %d DoF, nnz %d"
                   % (E, N, nnz), "path": path, "elements": E, "dofs": N,
                   "nnz": nnz},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "alg_bytes_per_launch": alg_bytes, "kernel": kernel_name},
        "dofs_per_s": N / (ms * 1e-3),
        "first_call_ms": first_call_ms,
        "end_to_end_ms": e2e_ms,
        "max_rel_err_vs_first_call": err,
        **extra,
    }
    # device-copy bandwidth (read + write bytes of a 1 GiB copy), for scale
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    cms = time_stream(lambda: dst.copy_(src), 10, stream)
    res["copy_bw_GBps"] = 2 * src.numel() * 8 / (cms * 1e-3) / 1e9
    res["roofline"]["frac_of_copy_bw"] = achieved / 1e9 / res["copy_bw_GBps"]
    del src, dst
    if not args.no_cpu:
        log('cpu baseline ...')
        res["cpu_baseline"] = cpu_baseline()
    print(json.dumps(res))


if __name__ == '__main__':
    main()
