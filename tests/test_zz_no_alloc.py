"""The allocation-free fused entry point (femhip_fused_assemble_static,
states and suspect queue in the library's static device array) against
femhip_fused_assemble with caller-allocated state and queue: bitwise equal
CSR values, load vector and verdict on two calls in a row, the second on a
smaller mesh; and the host-side rejection of a pattern with more exact
zeros than the static queue holds."""
import ctypes

import numpy as np
import pytest

import femhip
from femhip import _lib


def _fake(n):
    return ctypes.c_void_p(0x1000 * n)


def test_static_rejects_a_pattern_beyond_its_queue():
    lib = _lib.load()
    cap = lib.femhip_fused_static_queue()
    assert cap == 1 << 16
    info = _lib.FusedInfo()
    info.block = 512
    dev = _lib.FusedDev(None, None, None, None, None, info)
    chk = _lib.FusedCheck(_fake(1), _fake(2), _fake(3), 10, 4, 0, _fake(4), cap + 1,
                          None, 0, _fake(5))
    rc = lib.femhip_fused_assemble_static(ctypes.byref(dev), _lib.LAPLACE, 3, _fake(6), 10,
                                          None, _fake(7), None, ctypes.byref(chk), 0, None)
    assert rc == _lib.EINVAL and b"static suspect queue" in lib.femhip_last_error()
    rc = lib.femhip_fused_assemble_static(ctypes.byref(dev), 7, 3, _fake(6), 10, None,
                                          _fake(7), None, None, 0, None)
    assert rc == _lib.EUNSUPPORTED


def _prepared(n, seed):
    """A jittered n^3 P1-tet Laplace whose pattern has a fused plan."""
    from femhip.models import laplace
    x = np.linspace(0, 1, n)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1.0 / (n - 1), seed=seed)
    b = femhip.Basis(m, femhip.ElementTetP1())
    laplace.assemble_device(b)           # exact path: pattern validated
    laplace.assemble_device(b)           # fused plan in use
    d = b.device()
    plan = d.fused_last['laplace']
    return d, plan, d.fused_plans[plan.signature]


@pytest.mark.gpu
def test_static_entry_point_matches_the_allocating_one():
    import torch
    lib = _lib.load()
    cases = []
    for n, seed, with_load in ((12, 3, True), (7, 4, False)):   # the second smaller
        d, plan, fp = _prepared(n, seed)
        out = torch.full((plan.nnz,), float('nan'), dtype=torch.float64, device='cuda')
        load = (torch.full((fp.nnodes,), float('nan'), dtype=torch.float64, device='cuda')
                if with_load else None)
        verdict = torch.full((4,), -1, dtype=torch.int32, device='cuda')
        chk = _lib.FusedCheck.from_buffer_copy(fp.check_struct(d, plan.mask, plan.nzeros))
        chk.queue, chk.capacity = None, 0        # ignored by the static entry point
        chk.verdict = verdict.data_ptr()
        cases.append((n, d, plan, fp, out, load, verdict, chk))
    torch.cuda.synchronize()
    # two static calls in a row (the second on the smaller mesh)
    for n, d, plan, fp, out, load, verdict, chk in cases:
        print('static call, n =', n, 'nnz =', plan.nnz, flush=True)
        _lib.check(lib.femhip_fused_assemble_static(
            ctypes.byref(fp.dev), _lib.LAPLACE, 3, _lib.ptr(d.p), fp.nnodes, None,
            _lib.ptr(out), _lib.ptr(load), ctypes.byref(chk), 0, _lib.stream_ptr()))
    torch.cuda.synchronize()
    # the allocating entry point on the same inputs
    for n, d, plan, fp, out, load, verdict, chk in cases:
        print('reference call, n =', n, flush=True)
        ref = torch.empty_like(out)
        ref_load = torch.empty_like(load) if load is not None else None
        fp.assemble('laplace', d.p, ref, load=ref_load,
                    check=fp.check_struct(d, plan.mask, plan.nzeros))
        torch.cuda.synchronize()
        v_ref = fp.verdict.cpu().tolist()
        v = verdict.cpu().tolist()
        assert v_ref[0] == 1 and v == v_ref, (v, v_ref)
        assert torch.equal(out, ref), float((out - ref).abs().max())
        if load is not None:
            assert torch.equal(load, ref_load)
