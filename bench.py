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


def measured_traffic(kernel_name, n):
    """HBM bytes per launch of the headline kernel from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain
    copy), or (None, reason) when there is none for this kernel/workload."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                        '%s_tet%d_laplace.md' % (kernel_name.split(' ')[0], n))
    try:
        with open(path) as f:
            m = re.search(r'traffic_bytes_per_launch:\s*(\d+)', f.read())
    except OSError:
        return None, 'no counter profile for this kernel/workload'
    if not m:
        return None, 'no traffic line in ' + os.path.basename(path)
    return int(m.group(1)), 'profiles/' + os.path.basename(path)


def _median_s(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(n_sample=61, reps=3):
    """The oracle (NumPy/SciPy restatement of the reference algorithm) on a
    bounded sample of the headline workload: MeshTet 61^3 jittered
    (1,296,000 tets) -- Basis construction, laplace, mass and unit_load,
    each the median of ``reps`` runs -- with the reference's thread knob at
    the host's cores (nthreads, the (j, i) entries split over Python
    threads, bilinear_form.py:100-119) and with its default (nthreads=0)."""
    from oracle import skfem_np as O
    m = build_mesh(n_sample, 0.2)
    om = O.MeshTet(m.p, m.t)
    E = m.nelements
    cores = len(os.sched_getaffinity(0))
    holder = {}

    def basis():
        holder['b'] = O.CellBasis(om, O.ElementTetP1())
    t_basis = _median_s(basis, reps)
    ob = holder['b']
    out = {}
    for nth in (cores, 0):
        tl = _median_s(lambda: O.bilinear_assemble(O.form_laplace, ob, nthreads=nth), reps)
        tm = _median_s(lambda: O.bilinear_assemble(O.form_mass, ob, nthreads=nth), reps)
        tu = _median_s(lambda: O.linear_assemble(O.form_unit_load, ob), reps)
        out[nth] = (tl, tm, tu)
    tl, tm, tu = out[cores]
    tl0, tm0, tu0 = out[0]
    return {"value": E / (t_basis + tl + tm + tu), "unit": "elements/s", "cores": cores,
            "kind": "port",
            "sample": "oracle on MeshTet %d^3 jittered (%d tets): Basis %.2fs, laplace %.2fs, "
                      "mass %.2fs, unit_load %.2fs with nthreads=%d (median of %d; NumPy/SciPy)"
                      % (n_sample, E, t_basis, tl, tm, tu, cores, reps),
            "single_thread_value": E / (t_basis + tl0 + tm0 + tu0),
            "single_thread_s": {"basis": t_basis, "laplace": tl0, "mass": tm0, "unit_load": tu0},
            "host_cpus": os.cpu_count(), "host_affinity": cores}


def cpu_configs():
    """One-time timings of the oracle on every benchmark configuration (the
    reference's thread knob at the host's cores), progress on stderr:
    P1 triangles 129^2 (laplace + unit_load), P1 tetrahedra 101^3 jittered
    (laplace, mass with nthreads = cores and = 0, unit_load; each the median
    of 3), P2 tetrahedra 41^3 jittered,
    vector Q2 hexahedral elasticity 20^3 jittered (the 40^3 case is 8x the
    elements), kappa-Laplace on P1 tetrahedra 65^3, one kappa sample on P2
    triangles 257^2 and the 64 P1 triangle meshes (per-mesh p50/p99)."""
    from oracle import skfem_np as O
    cores = len(os.sched_getaffinity(0))
    res = {"cores": cores}

    def tick(name):
        def f(j, i):
            if i == 0:
                log('  %s: test functions of trial %d done' % (name, j))
        return f

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return time.perf_counter() - t0, r
    # config 1
    x = np.linspace(0, 1, 129)
    m = O.MeshTri.init_tensor(x, x)
    tb, b = timed(lambda: O.CellBasis(m, O.ElementTriP1()))
    tl, _ = timed(lambda: O.bilinear_assemble(O.form_laplace, b, nthreads=cores))
    tu, _ = timed(lambda: O.linear_assemble(O.form_unit_load, b))
    res["tri1_n129"] = {"elements": b.nelems, "basis_s": tb, "laplace_s": tl, "unit_load_s": tu}
    log('cpu config tri1_n129 done')
    # config 2a
    mm = build_mesh(101, 0.2)
    m = O.MeshTet(mm.p, mm.t)
    tb, b = timed(lambda: O.CellBasis(m, O.ElementTetP1()))
    r = {"elements": b.nelems, "basis_s": tb}
    for nth, tag in ((cores, ''), (0, '_nthreads0')):
        for name, f in (('laplace', O.form_laplace), ('mass', O.form_mass)):
            r[name + tag + '_s'] = _median_s(
                lambda: O.bilinear_assemble(f, b, nthreads=nth), 3)
            log('cpu config tet1_n101j', name, 'nthreads', nth, 'done')
    r['unit_load_s'] = _median_s(lambda: O.linear_assemble(O.form_unit_load, b), 3)
    res["tet1_n101j"] = r
    del b
    # config 2b
    x = np.linspace(0, 1, 41)
    m = O.jitter(O.MeshTet.init_tensor(x, x, x), 0.2, 1 / 40)
    tb, b = timed(lambda: O.CellBasis(m, O.ElementTetP2(), intorder=4))
    tl, _ = timed(lambda: O.bilinear_assemble(O.form_laplace, b, nthreads=cores,
                                              progress=tick('tet2_n41j')))
    res["tet2_n41j"] = {"elements": b.nelems, "basis_s": tb, "laplace_s": tl}
    log('cpu config tet2_n41j done')
    del b
    # config 3 at 20^3
    x = np.linspace(0, 1, 21)
    m = O.jitter(O.MeshHex.init_tensor(x, x, x), 0.15, 1 / 20)
    lam, mu = O.lame_parameters(1.0, 0.3)
    tb, b = timed(lambda: O.CellBasis(m, O.ElementVector(O.ElementHex2()), intorder=5))
    tl, _ = timed(lambda: O.bilinear_assemble(O.form_elasticity(lam, mu), b, nthreads=cores,
                                              progress=tick('vhex2_n21j')))
    res["vhex2_n21j"] = {"elements": b.nelems, "basis_s": tb, "elasticity_s": tl}
    log('cpu config vhex2_n21j done')
    del b
    # config 4
    x = np.linspace(0, 1, 65)
    m = O.MeshTet.init_tensor(x, x, x)
    kap = np.random.default_rng(0).lognormal(0, 1, m.t.shape[1])
    tb, b = timed(lambda: O.CellBasis(m, O.ElementTetP1()))
    kq = np.repeat(kap[:, None], b.dx.shape[1], axis=1)
    tl, _ = timed(lambda: O.bilinear_assemble(O.form_kappa_laplace, b, nthreads=cores, kappa=kq))
    res["tet1_n65_kappa"] = {"elements": b.nelems, "basis_s": tb, "kappa_laplace_s": tl}
    log('cpu config tet1_n65 done')
    del b
    # config 5a: one sample
    x = np.linspace(0, 1, 257)
    m = O.MeshTri.init_tensor(x, x)
    tb, b = timed(lambda: O.CellBasis(m, O.ElementTriP2()))
    kap = np.random.default_rng(0).lognormal(0, 0.5, m.t.shape[1])
    kq = np.repeat(kap[:, None], b.dx.shape[1], axis=1)
    tl, _ = timed(lambda: O.bilinear_assemble(O.form_kappa_laplace, b, nthreads=cores, kappa=kq))
    res["tri2_n257_sample"] = {"elements": b.nelems, "basis_s": tb, "kappa_laplace_s": tl,
                               "x256_s": 256 * tl}
    log('cpu config tri2_n257 done')
    del b
    # config 5b
    per = []
    for n in np.random.default_rng(0).integers(17, 258, 64):
        x = np.linspace(0, 1, int(n))
        m = O.MeshTri.init_tensor(x, x)
        t0 = time.perf_counter()
        b = O.CellBasis(m, O.ElementTriP1())
        O.bilinear_assemble(O.form_laplace, b)
        O.linear_assemble(O.form_unit_load, b)
        per.append(time.perf_counter() - t0)
    res["mesh_set_64"] = {"total_s": float(np.sum(per)), "p50_s": _pct(per, 50),
                          "p99_s": _pct(per, 99)}
    log('cpu config mesh_set_64 done')
    return res


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def bench_batched_kappa(S=256, n=257):
    """256 per-element kappa samples on the P2 triangle mesh 257^2 in one
    [S, nnz] launch vs plain per-sample GPU assemblies."""
    import torch
    import femhip
    from femhip.batch import KappaSamples
    from femhip.models import kappa_laplace
    x = np.linspace(0, 1, n)
    m = femhip.MeshTri.init_tensor(x, x)
    b = femhip.Basis(m, femhip.ElementTriP2())
    kap = np.random.default_rng(0).lognormal(0, 0.5, (S, m.nelements))
    kt = torch.tensor(kap, device='cuda')
    ks = KappaSamples(b)
    out = torch.empty((S, ks.nnz), dtype=torch.float64, device='cuda')
    stream = torch.cuda.current_stream()
    ks.assemble(kt, out)
    ms_batch = time_stream(lambda: ks.assemble(kt, out), 5, stream)
    # the same launch captured in a HIP graph and replayed
    kg = kt.clone()
    og = torch.empty_like(out)
    ks.capture(kg, og)
    ms_graph = time_stream(ks.replay, 5, stream)
    torch.cuda.synchronize()
    assert torch.equal(og, out)
    # plain: one drop-in GPU assembly per sample (exact path; kappa changes
    # every call, so each call also validates its pattern)
    nplain = 8
    kappa_laplace.assemble_device(b, kappa=kap[0][:, None])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(1, nplain + 1):
        kappa_laplace.assemble_device(b, kappa=kt[s][:, None])
    torch.cuda.synchronize()
    ms_plain = (time.perf_counter() - t0) * 1e3 / nplain
    return {"samples": S, "elements": m.nelements, "dofs": b.N, "nnz": ks.nnz,
            "batched_ms_total": ms_batch, "batched_ms_per_sample": ms_batch / S,
            "graph_replay_ms_total": ms_graph,
            "plain_ms_per_sample": ms_plain,
            "speedup_per_sample": ms_plain / (ms_batch / S),
            "batched_GBps_written": S * ks.nnz * 8 / (ms_batch * 1e-3) / 1e9}


def bench_mesh_batch(k=64):
    """64 P1 triangle meshes, n_i ~ U[17, 257]: Laplace + unit load, one
    persistent graph-captured launch vs per-mesh launches."""
    import torch
    import femhip
    from femhip.batch import MeshBatch
    from femhip.models import laplace
    ns = np.random.default_rng(0).integers(17, 258, k)
    bases = []
    for n in ns:
        x = np.linspace(0, 1, int(n))
        bases.append(femhip.Basis(femhip.MeshTri.init_tensor(x, x), femhip.ElementTriP1()))
    mb = MeshBatch(bases, groups=512)
    mb.capture()
    stream = torch.cuda.current_stream()
    ms_graph = time_stream(mb.replay, 20, stream)
    ms_launch = time_stream(mb.assemble, 20, stream)
    # plain: per-mesh fused launches (Laplace + load), each timed alone
    per = []
    for b in bases:
        d = b.device()
        plan = next(p for kk, p in d.validated.items() if kk[0] == 'laplace')
        fp = d.fused_plans[plan.signature]
        out = torch.empty(plan.nnz, dtype=torch.float64, device='cuda')
        ld = torch.empty(b.N, dtype=torch.float64, device='cuda')
        per.append(time_stream(lambda: fp.assemble('laplace', d.p, out, load=ld), 10, stream))
    E = sum(b.nelems for b in bases)
    return {"meshes": k, "elements": int(E), "graph_ms": ms_graph, "launch_ms": ms_launch,
            "per_mesh_batched_ms": ms_graph / k, "plain_sum_ms": float(np.sum(per)),
            "plain_p50_ms": _pct(per, 50), "plain_p99_ms": _pct(per, 99),
            "speedup_vs_plain_sum": float(np.sum(per)) / ms_graph}


def bench_adjoint(n=65):
    """kappa-Laplace + load on MeshTet 65^3: forward, the adjoint alone
    (gradients of given upstream dL/dA, dL/db w.r.t. kappa and the node
    coordinates), and a full forward + loss + backward step."""
    import torch
    import femhip
    from femhip.autograd import kappa_laplace_with_load, quadratic_loss
    x = np.linspace(0, 1, n)
    m = femhip.MeshTet.init_tensor(x, x, x)
    b = femhip.Basis(m, femhip.ElementTetP1())
    kap = torch.tensor(np.random.default_rng(0).lognormal(0, 1, m.nelements), device='cuda')
    u = torch.tensor(np.random.default_rng(1).standard_normal(b.N), device='cuda')
    f = torch.tensor(np.random.default_rng(2).standard_normal(b.N), device='cuda')
    p = torch.tensor(m.p, device='cuda')
    stream = torch.cuda.current_stream()
    k = kap.clone().requires_grad_(True)
    pp = p.clone().requires_grad_(True)
    data, load, (ip, ix) = kappa_laplace_with_load(b, k, pp)
    gd = torch.randn_like(data)
    gl = torch.randn_like(load)

    def fwd():
        with torch.no_grad():
            kappa_laplace_with_load(b, kap, p)

    def bwd():
        torch.autograd.grad((data, load), (k, pp), grad_outputs=(gd, gl), retain_graph=True)

    def fwdbwd():
        k2 = kap.clone().requires_grad_(True)
        p2 = p.clone().requires_grad_(True)
        d2, l2, _ = kappa_laplace_with_load(b, k2, p2)
        quadratic_loss(d2, ip, ix, u, f).backward()
    fwdbwd()
    bwd()
    ms_f = time_stream(fwd, 10, stream)
    ms_b = time_stream(bwd, 10, stream)
    ms_fb = time_stream(fwdbwd, 10, stream)
    # the launch sequences alone (C ABI, no autograd bookkeeping): the
    # forward's fused pass and the fused adjoint
    import ctypes
    from femhip import _lib
    from femhip.autograd import _ctx
    c = _ctx(b)
    fo = torch.empty_like(data)
    fl = torch.empty_like(load)
    dk = torch.empty_like(kap)
    dX = torch.empty_like(p)

    def fwd_launch():
        c.fp.assemble('kappa_laplace', p, fo, kap, load=fl)

    def adj_launch():
        _lib.check(c.d.lib.femhip_fused_adjoint(
            ctypes.byref(c.fp.dev), _lib.ptr(c.fp.state('adjoint')), 3, _lib.ptr(p), p.shape[1],
            _lib.ptr(c.plan.indptr), _lib.ptr(c.diag_slot), _lib.ptr(kap), _lib.ptr(gd),
            _lib.ptr(gl), _lib.ptr(dk), _lib.ptr(dX), _lib.stream_ptr()))
    fwd_launch()
    adj_launch()
    ms_fl = time_stream(fwd_launch, 20, stream)
    ms_al = time_stream(adj_launch, 20, stream)
    # algorithmic bytes: connectivity, coordinates and kappa read, values and
    # load written (forward); + dL/dA, dL/db read, dkappa and dX written
    # (adjoint; dL/dA on the CSR pattern)
    E, N, nnz = m.nelements, b.N, int(data.numel())
    fwd_bytes = 16 * E + 24 * N + 8 * E + 8 * nnz + 8 * N
    bwd_bytes = 16 * E + 24 * N + 8 * E + 8 * nnz + 8 * N + 8 * E + 24 * N
    return {"elements": E, "dofs": N, "nnz": nnz,
            "forward_ms": ms_f, "adjoint_ms": ms_b, "adjoint_over_forward": ms_b / ms_f,
            "forward_loss_backward_ms": ms_fb,
            "forward_frac_of_8TBps": fwd_bytes / (ms_f * 1e-3) / HBM_PEAK,
            "adjoint_alg_bytes": bwd_bytes,
            "adjoint_frac_of_8TBps": bwd_bytes / (ms_b * 1e-3) / HBM_PEAK,
            "forward_launch_ms": ms_fl, "adjoint_launch_ms": ms_al,
            "adjoint_launch_over_forward_launch": ms_al / ms_fl,
            "adjoint_launch_frac_of_8TBps": bwd_bytes / (ms_al * 1e-3) / HBM_PEAK}


def bench_changed_inputs(basis, calls=10):
    """The drop-in ``assemble_device`` on the 101^3 basis with a new input on
    every call, each call validated in its own launch sequence (fused pass,
    recheck of the entries within the rounding bound of zero, verdict read
    back): a new per-element kappa ~ lognormal(0, 1) (kappa-Laplace), and
    new interior-node coordinates (jitter U(-0.2h, 0.2h), a new seed per
    call; Laplace).  Wall time per call including the verdict's
    synchronisation; the fraction of calls that stayed on the fused path."""
    import torch
    from femhip.models import laplace, kappa_laplace
    d = basis.device()
    m = basis.mesh
    E = m.nelements
    rng = np.random.default_rng(21)
    kappas = [torch.tensor(rng.lognormal(0, 1, E), device='cuda')[:, None]
              for _ in range(calls + 1)]
    x = np.linspace(0, 1, int(round(m.p.shape[1] ** (1 / 3))))
    base = type(m).init_tensor(x, x, x)
    I = base.interior_nodes()
    h = 1.0 / (len(x) - 1)
    coords = []
    for _ in range(calls + 1):
        p = base.p.copy()
        p[:, I] += rng.uniform(-0.2 * h, 0.2 * h, size=(3, len(I)))
        coords.append(torch.tensor(p, device='cuda'))
    p0 = d.p
    res = {}
    for name, run in (('new_kappa', lambda k: kappa_laplace.assemble_device(basis, kappa=kappas[k])),
                      ('new_coordinates', None)):
        if name == 'new_coordinates':
            def run(k):
                d.set_coordinates(coords[k])
                return laplace.assemble_device(basis)
        run(calls)                       # first call of the form: exact path
        run(calls)                       # the fused path's first call (its exponent)
        torch.cuda.synchronize()
        fused = 0
        t0 = time.perf_counter()
        for k in range(calls):
            run(k)
            fused += int(d.last_verdict is not None and d.last_verdict[0] == 1)
        torch.cuda.synchronize()
        res[name + '_ms'] = (time.perf_counter() - t0) * 1e3 / calls
        res[name + '_fused_fraction'] = fused / calls
    # the same inputs through the validated launch sequence alone (fused
    # pass, repeat launch, recheck, verdict; HIP events, no host
    # synchronisation per call), as the headline is timed: a new input on
    # every launch
    stream = torch.cuda.current_stream()
    kflat = [kp[:, 0].contiguous() for kp in kappas]
    for name, form, inputs in (('new_kappa', 'kappa_laplace', kflat),
                               ('new_coordinates', 'laplace', coords)):
        plan = d.fused_last[form]
        fp = d.fused_plans[plan.signature]
        chk = fp.check_struct(d, plan.mask, plan.nzeros)
        out = torch.empty(plan.nnz, dtype=torch.float64, device='cuda')
        k = [0]

        def step():
            i = k[0] % calls
            k[0] += 1
            if form == 'kappa_laplace':
                fp.assemble(form, p0, out, kappa=inputs[i], check=chk, groups=fp.groups)
            else:
                fp.assemble(form, inputs[i], out, check=chk, groups=fp.groups)
        step()
        torch.cuda.synchronize()
        res[name + '_sequence_ms'] = time_stream(step, 2 * calls, stream)
        res[name + '_sequence_last_verdict'] = fp.verdict.cpu().tolist()
    d.set_coordinates(p0)
    laplace.assemble_device(basis)
    return res


def bench_device_solve(basis, A):
    """The assembled 101^3 Laplace kept on the device: condense (Dirichlet
    boundary) + Jacobi-PCG to rtol 1e-8 (second calls, warm), beside the
    host hand-off a SciPy solve would start with (D2H copy + csr_matrix)."""
    import torch
    from femhip.models import unit_load
    from femhip.utils import condense_device, solve_device
    f = unit_load.assemble(basis)
    D = basis.mesh.boundary_nodes()
    # first calls (untimed): kernel loading and first allocations
    AII, bI, xx, I = condense_device(A, f, D=D)
    solve_device(AII, bI, xx, I, rtol=1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    AII, bI, xx, I = condense_device(A, f, D=D)
    torch.cuda.synchronize()
    t_cond = (time.perf_counter() - t0) * 1e3
    info = {}
    t0 = time.perf_counter()
    u = solve_device(AII, bI, xx, I, rtol=1e-8, info=info)
    torch.cuda.synchronize()
    t_solve = (time.perf_counter() - t0) * 1e3
    from scipy.sparse import csr_matrix
    t0 = time.perf_counter()
    _ = csr_matrix((A.data.cpu().numpy(), A.indices.cpu().numpy(), A.indptr.cpu().numpy()),
                   shape=A.shape)
    t_host = (time.perf_counter() - t0) * 1e3
    return {"dofs": int(A.shape[0]), "interior": int(AII.shape[0]), "condense_ms": t_cond,
            "pcg_ms": t_solve, "pcg_iterations": info["iterations"],
            "pcg_ms_per_iteration": t_solve / max(1, info["iterations"]),
            "final_residual": info["residual"], "umax": float(u.max()),
            "host_handoff_ms": t_host}


def bench_interpolate(basis):
    """Basis.interpolate of a DOF vector (values and gradients at the 4
    quadrature points) on the 101^3 P1-tet basis; bytes = DOF map + w +
    mapping read, values + gradients written."""
    import torch
    d = basis.device()
    w = torch.randn(basis.N, dtype=torch.float64, device='cuda')
    d.geometry()
    basis.interpolate_device(w)
    ms = time_stream(lambda: basis.interpolate_device(w), 20, torch.cuda.current_stream())
    E, Q, D = d.E, d.Q, d.dim
    nbytes = E * (4 * d.nbfun + 8 * (D * D + 1)) + 8 * basis.N + 8 * E * Q * (1 + D)
    return {"elements": E, "points": Q, "ms": ms, "GBps": nbytes / (ms * 1e-3) / 1e9,
            "frac_of_8TBps": nbytes / (ms * 1e-3) / HBM_PEAK}


def bench_facet(m):
    """Boundary forms on the boundary facets of the 101^3 jittered tet mesh
    with P1 (FacetBasis): facet mass (local entries + segmented sums over the
    cached pattern) and facet load.  Local-kernel bytes per facet: facet
    vertices, tind and t[0, e] (4 B each), their coordinates (8 B x dim x
    (dim + 1)), the element mapping rows (8 B x dim^2) read, nbfun^2 doubles
    written."""
    import torch
    import femhip
    from femhip.models import mass, unit_load
    t0 = time.perf_counter()
    fb = femhip.FacetBasis(m, femhip.ElementTetP1())
    d = fb.device()
    A = mass.assemble_device(fb)
    b_first = unit_load.assemble_device(fb)
    torch.cuda.synchronize()
    first = (time.perf_counter() - t0) * 1e3
    plan = next(iter(d.plans.values()))
    nf, nb, D = d.E, d.nbfun, d.dim
    stream = torch.cuda.current_stream()
    contrib = torch.empty(nb * nb * nf, dtype=torch.float64, device='cuda')
    out = torch.empty(max(plan.nnz, 1), dtype=torch.float64, device='cuda')[:plan.nnz]

    def step():
        d.local_bilinear('mass', out=contrib)
        d.segment_sum(contrib, plan.perm, plan.seg, plan.nnz, out=out)
    step()
    ms = time_stream(step, 20, stream)
    local_ms = time_stream(lambda: d.local_bilinear('mass', out=contrib), 20, stream)
    perm, seg = d.rows_plan()
    bout = torch.empty(fb.N, dtype=torch.float64, device='cuda')

    def lstep():
        d.segment_sum(d.local_linear('unit_load'), perm, seg, fb.N, out=bout)
    lstep()
    load_ms = time_stream(lstep, 20, stream)
    torch.cuda.synchronize()
    err = float((out - A.data).abs().max() / A.data.abs().max())
    errb = float((bout - b_first).abs().max() / b_first.abs().max())
    local_bytes = nf * (4 * (D + 2) + 8 * D * (D + 1) + 8 * D * D + 8 * nb * nb)
    return {"facets": nf, "nnz": plan.nnz, "first_call_ms": first,
            "mass_ms": ms, "mass_facets_per_s": nf / (ms * 1e-3),
            "local_kernel_ms": local_ms,
            "local_kernel_GBps": local_bytes / (local_ms * 1e-3) / 1e9,
            "load_ms": load_ms, "max_rel_err_vs_first_call": max(err, errb)}


def bench_p2_tet(n=41):
    """Laplace on ElementTetP2, MeshTet 41^3 jittered (384,000 tets,
    531,441 DoF, 11 points at intorder 4), later calls: the fused rows
    kernel (entries in the reference's arithmetic summed into LDS rows, the
    pattern checked in the same pass, verdict read back by the drop-in
    call), beside the exact path it replaces (local entries + segmented
    sums over the cached pattern).  Algorithmic bytes: 40 B element DOFs per
    element + 24 B coordinates per mesh vertex + 8 B per CSR value."""
    import torch
    import femhip
    from femhip.models import laplace
    x = np.linspace(0, 1, n)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1.0 / (n - 1))
    b = femhip.Basis(m, femhip.ElementTetP2(), intorder=4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A = laplace.assemble_device(b)
    torch.cuda.synchronize()
    first = (time.perf_counter() - t0) * 1e3
    d = b.device()
    plan = d.rows_last['laplace']
    fp = d.rows_plans[plan.signature]
    E, nb = d.E, d.nbfun
    stream = torch.cuda.current_stream()
    out = torch.empty(plan.nnz, dtype=torch.float64, device='cuda')
    fp.assemble('laplace', d, out)
    fused_ms = time_stream(lambda: fp.assemble('laplace', d, out), 10, stream)
    torch.cuda.synchronize()
    verdict = fp.verdict.cpu().tolist()
    err = float((out - A.data).abs().max() / A.data.abs().max())
    # the exact path's steady state on the same pattern
    contrib = torch.empty(nb * nb * E, dtype=torch.float64, device='cuda')
    out2 = torch.empty(plan.nnz, dtype=torch.float64, device='cuda')
    d.local_bilinear('laplace', out=contrib)
    local_ms = time_stream(lambda: d.local_bilinear('laplace', out=contrib), 10, stream)
    sum_ms = time_stream(lambda: d.segment_sum(contrib, plan.perm, plan.seg, plan.nnz,
                                               out=out2), 10, stream)
    torch.cuda.synchronize()
    A2 = laplace.assemble_device(b)          # the drop-in call (fused, validated)
    torch.cuda.synchronize()
    dropin = bool(d.last_verdict and d.last_verdict[0] == 1 and A2.nnz == A.nnz)
    # new interior coordinates on every drop-in call (validated in the pass)
    rng = np.random.default_rng(31)
    I = femhip.MeshTet.init_tensor(x, x, x).interior_nodes()
    base = femhip.MeshTet.init_tensor(x, x, x).p
    coords = []
    for _ in range(4):
        p = base.copy()
        p[:, I] += rng.uniform(-0.2 / (n - 1), 0.2 / (n - 1), size=(3, len(I)))
        coords.append(torch.tensor(p, device='cuda'))
    p0 = d.p
    fused_calls = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for pc in coords:
        d.set_coordinates(pc)
        laplace.assemble_device(b)
        fused_calls += int(d.last_verdict is not None and d.last_verdict[0] == 1)
    torch.cuda.synchronize()
    new_coords_ms = (time.perf_counter() - t0) * 1e3 / len(coords)
    d.set_coordinates(p0)
    alg = 40 * E + 24 * m.p.shape[1] + 8 * plan.nnz
    return {"elements": E, "dofs": b.N, "nnz": plan.nnz, "points": d.Q,
            "first_call_ms": first, "fused_rows_ms": fused_ms,
            "elements_per_s": E / (fused_ms * 1e-3),
            "alg_bytes": alg, "frac_of_8TBps": alg / (fused_ms * 1e-3) / HBM_PEAK,
            "verdict": verdict, "fused_plan": fp.stats(),
            "max_rel_err_vs_first_call": err,
            "dropin_fused": dropin, "new_coordinates_ms": new_coords_ms,
            "new_coordinates_fused_fraction": fused_calls / len(coords),
            "exact_path_local_ms": local_ms, "exact_path_segment_sum_ms": sum_ms,
            "exact_path_ms": local_ms + sum_ms}


def bench_laplace_mass(basis):
    """Config 2a's workload on the 101^3 basis: laplace, mass and unit_load
    in one fused pass (checked against the Laplace pattern), beside the
    separate validated fused calls.  Algorithmic bytes: 16 B connectivity
    per element + 24 B coordinates per node + 8 B per value of both CSR
    matrices + 8 B per load entry."""
    import torch
    from femhip.fused import LaplaceMassPass
    from femhip.models import unit_load
    d = basis.device()
    pair = LaplaceMassPass(basis)
    stream = torch.cuda.current_stream()
    L, M, f = pair.assemble()
    ms_pair = time_stream(lambda: pair.assemble(L, M, f), 20, stream)
    torch.cuda.synchronize()
    verdict = pair.fp.verdict.cpu().tolist()
    # the separate passes (fused, validated) and the load
    pl, pm = pair.plan_l, pair.plan_m
    fl, fm = d.fused_plans[pl.signature], d.fused_plans[pm.signature]
    cl = fl.check_struct(d, pl.mask, pl.nzeros)
    cm = fm.check_struct(d, pm.mask, pm.nzeros)
    ol = torch.empty_like(L)
    om = torch.empty_like(M)
    ms_l = time_stream(lambda: fl.assemble('laplace', d.p, ol, check=cl), 20, stream)
    ms_m = time_stream(lambda: fm.assemble('mass', d.p, om, check=cm), 20, stream)
    ms_f = time_stream(lambda: unit_load.assemble_device(basis), 5, stream)
    torch.cuda.synchronize()
    err_l = float((L - ol).abs().max() / ol.abs().max())
    err_m = float((M - om).abs().max() / om.abs().max())
    E, N = d.E, d.N
    alg = 16 * E + 24 * N + 8 * pl.nnz + 8 * pm.nnz + 8 * N
    return {"elements": E, "nnz_laplace": pl.nnz, "nnz_mass": pm.nnz,
            "pair_ms": ms_pair, "alg_bytes": alg, "alg_bytes_per_element": alg / E,
            "frac_of_8TBps": alg / (ms_pair * 1e-3) / HBM_PEAK,
            "elements_per_s": E / (ms_pair * 1e-3), "verdict": verdict,
            "separate_laplace_ms": ms_l, "separate_mass_ms": ms_m,
            "separate_unit_load_ms": ms_f, "separate_total_ms": ms_l + ms_m + ms_f,
            "max_rel_diff_vs_separate": max(err_l, err_m),
            "pair_plan": pair.fp.stats()}


def bench_autotune(basis):
    """The fused kernel's autotuner on the 101^3 Laplace: every candidate
    launch configuration timed, the fastest installed on the basis (the
    headline above uses the default configuration)."""
    from femhip import autotune
    rec = autotune.tune_fused(basis, 'laplace', steps=10)
    default = next(ms for cfg, ms, dup in rec['trials']
                   if tuple(cfg) == tuple(autotune.CANDIDATES[0]))
    return {"best": rec['best'], "best_ms": rec['ms'], "default_ms": default,
            "trials": [[cfg, round(ms, 4), round(dup, 3)] for cfg, ms, dup in rec['trials']]}


def bench_elasticity(n=41):
    """linear_elasticity on ElementVector(ElementHex2), 27 points, MeshHex
    (n-1)^3 jittered 0.15h: MFMA local matrices + segmented sums."""
    import torch
    import femhip
    from femhip.models import linear_elasticity, lame_parameters
    x = np.linspace(0, 1, n)
    m = femhip.MeshHex.init_tensor(x, x, x).jittered(0.15, 1.0 / (n - 1))
    b = femhip.Basis(m, femhip.ElementVector(femhip.ElementHex2()), intorder=5)
    lam, mu = lame_parameters(1.0, 0.3)
    form = linear_elasticity(lam, mu)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A = form.assemble_device(b)
    torch.cuda.synchronize()
    first = (time.perf_counter() - t0) * 1e3
    d = b.device()
    plan = next(iter(d.validated.values()))
    out = torch.empty(plan.nnz, dtype=torch.float64, device='cuda')
    loc = torch.empty(81 * 81 * m.nelements, dtype=torch.float64, device='cuda')
    G = d.geometry()
    stream = torch.cuda.current_stream()
    from femhip import _lib
    import ctypes

    def local():
        _lib.check(d.lib.femhip_hex2_elasticity_mfma(
            ctypes.byref(d.geom), ctypes.byref(d.space), lam, mu, _lib.ptr(G),
            _lib.ptr(loc), 0, _lib.stream_ptr()))
    local()     # one untimed launch: the first after the exact path runs slower
    ms_local = time_stream(local, 10, stream)
    # the first numeric call also builds the row-sum map of the pattern (once)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.elasticity_mfma(lam, mu, plan, out, loc)
    torch.cuda.synchronize()
    first_numeric = (time.perf_counter() - t0) * 1e3
    ms_all = time_stream(lambda: d.elasticity_mfma(lam, mu, plan, out, loc), 5, stream)
    err = float((out - A.data).abs().max() / A.data.abs().max())
    ndiff = int((out != A.data).sum())    # the two orders round differently
    E = m.nelements
    exec_flops = E * 6 * 2 * 32 * 32 * 28
    return {"elements": E, "dofs": b.N, "nnz": plan.nnz, "first_call_ms": first,
            "first_numeric_call_ms": first_numeric,
            "mfma_local_ms": ms_local, "assembly_ms": ms_all,
            "segment_sum_ms": ms_all - ms_local,
            "mfma_TFLOPs_executed": exec_flops / (ms_local * 1e-3) / 1e12,
            "mfma_util_of_78.6TF": exec_flops / (ms_local * 1e-3) / 78.6e12,
            "max_rel_err_vs_exact": err, "entries_differing_from_exact": ndiff}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--n', type=int, default=101)
    ap.add_argument('--path', default='auto', choices=['auto', 'generic', 'fused'])
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-configs', action='store_true',
                    help='time the oracle on every benchmark configuration (minutes; '
                         'prints its own JSON line, no GPU work)')
    ap.add_argument('--no-extra', action='store_true',
                    help='skip the batched / adjoint / elasticity measurements')
    ap.add_argument('--extra', default=None,
                    help='comma-separated subset of the extra measurements '
                         '(changed, interpolate, facet, device_solve, batched_kappa, '
                         'mesh_batch, adjoint, elasticity, p2_tet, laplace_mass, '
                         'autotune)')
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.cpu_configs:
        print(json.dumps({"cpu_configs": cpu_configs()}))
        return

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
        plan = dvb.fused_last['laplace']
        fp = dvb.fused_plans[plan.signature]
        extra['fused_plan'] = fp.stats()
        # the launch sequence of a drop-in call: the fused pass, the repeat
        # launch (a pass only when the exponent changed), the recheck of the
        # entries within the rounding bound of zero and the verdict
        chk = fp.check_struct(dvb, plan.mask, plan.nzeros)

        def step():
            fp.assemble('laplace', dvb.p, out, check=chk, groups=fp.groups)

        def step_unchecked():
            fp.assemble('laplace', dvb.p, out, groups=fp.groups)
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
    if path == 'fused':
        verdict = fp.verdict.cpu().tolist()
        assert verdict[0] == 1, verdict
        extra['fused_verdict'] = verdict
        # the bare pass (fused + repeat launch) without the validation
        extra['fused_unchecked_ms'] = time_stream(step_unchecked, args.steps, stream)

    # end-to-end drop-in call: BilinearForm.assemble returning a SciPy CSR
    # matrix (numeric assembly + D2H of the values + csr_matrix), steady state
    laplace.assemble(basis)
    t0 = time.perf_counter()
    for _ in range(3):
        laplace.assemble(basis)
    e2e_ms = (time.perf_counter() - t0) * 1e3 / 3

    alg_bytes = 16 * E + 24 * Nv + 8 * nnz
    achieved = alg_bytes / (ms * 1e-3)
    traffic, traffic_src = measured_traffic(kernel_name, args.n)
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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": traffic_src,
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
    if not args.no_extra:
        extras = (
            ('changed', "changed_inputs_tet101", lambda: bench_changed_inputs(basis)),
            ('interpolate', "interpolate_tet101", lambda: bench_interpolate(basis)),
            ('facet', "facet_mass_tet101_boundary", lambda: bench_facet(m)),
            ('device_solve', "device_solve_tet101", lambda: bench_device_solve(basis, A)),
            ('batched_kappa', "batched_kappa_p2tri257_S256", bench_batched_kappa),
            ('mesh_batch', "mesh_batch_64_p1tri", bench_mesh_batch),
            ('adjoint', "adjoint_kappa_tet65", bench_adjoint),
            ('elasticity', "elasticity_vhex2_40", bench_elasticity),
            ('p2_tet', "laplace_p2tet41", bench_p2_tet),
            ('laplace_mass', "laplace_mass_load_tet101", lambda: bench_laplace_mass(basis)),
            ('autotune', "autotune_tet101", lambda: bench_autotune(basis)),
        )
        chosen = None if args.extra is None else set(args.extra.split(','))
        for short, name, fn in extras:
            if chosen is not None and short not in chosen:
                continue
            log('extra:', name)
            try:
                res[name] = fn()
            except Exception as exc:   # report, keep the headline line intact
                res[name] = {"error": "%s: %s" % (type(exc).__name__, exc)}
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
    ch = res.get("changed_inputs_tet101")
    if isinstance(ch, dict) and "new_kappa_ms" in ch:
        ch["new_kappa_over_headline"] = ch["new_kappa_ms"] / ms
        ch["new_coordinates_over_headline"] = ch["new_coordinates_ms"] / ms
        for key in ("new_kappa", "new_coordinates"):
            if key + "_sequence_ms" in ch:
                ch[key + "_sequence_over_headline"] = ch[key + "_sequence_ms"] / ms
    if not args.no_cpu:
        log('cpu baseline ...')
        res["cpu_baseline"] = cpu_baseline()
    print(json.dumps(res))


if __name__ == '__main__':
    main()
