"""Post-assembly Dirichlet condition (``skfem/utils.py:295-400``)."""
import numpy as np


def enforce(A, b=None, x=None, I=None, D=None, diag=1., overwrite=False):
    """Zero the rows ``D`` of ``A`` (pattern kept), put ``diag`` on their
    diagonal and ``x[D]`` into ``b[D]``; returns ``A`` or ``(A, b)``."""
    if (I is None) == (D is None):
        raise Exception("Give only I or only D!" if I is not None
                        else "Either I or D must be given!")
    n = A.shape[0]
    if D is None:
        D = np.setdiff1d(np.arange(n, dtype=np.int32), np.asarray(I))
    D = np.asarray(D)
    if x is None:
        x = np.zeros(n, dtype=A.dtype)
    elif b is None:
        b = np.zeros_like(x)
    Aout = A if overwrite else A.copy()
    start = Aout.indptr[D]
    stop = Aout.indptr[D + 1]
    rows = np.repeat(np.arange(len(D)), stop - start)
    idx = np.repeat(start, stop - start) + (
        np.arange(rows.size) - np.repeat(np.cumsum(stop - start) - (stop - start),
                                         stop - start))
    Aout.data[idx] = 0.
    d = Aout.diagonal()
    d[D] = diag
    Aout.setdiag(d)
    if b is None:
        return Aout
    if hasattr(b, 'tocsr'):
        return Aout, enforce(b, D=D, diag=0., overwrite=overwrite)
    bout = b if overwrite else b.copy()
    bout[D] = x[D]
    return Aout, bout


def enforce_device(A, b=None, x=None, D=None, diag=1.):
    """``enforce`` on a ``DeviceCSR`` in place (rows of D zeroed, diagonal
    set, b[D] = x[D]); torch tensors stay on the device.  Returns (A, b)."""
    import torch
    from . import _lib
    lib = _lib.load()
    dev = A.data.device
    Dt = torch.as_tensor(np.unique(np.asarray(D)), dtype=torch.int32, device=dev)
    missing = torch.zeros(1, dtype=torch.int32, device=dev)
    xt = None if x is None else torch.as_tensor(x, dtype=torch.float64, device=dev)
    _lib.check(lib.femhip_enforce_rows(
        _lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(A.data), _lib.ptr(Dt), Dt.numel(),
        float(diag), _lib.ptr(b), _lib.ptr(xt), _lib.ptr(missing), _lib.stream_ptr()))
    if int(missing.item()):
        raise ValueError("a Dirichlet row has no diagonal slot; use femhip.enforce on the "
                         "SciPy matrix (it inserts one)")
    return A, b


def _dofs_split(n, I=None, D=None):
    """(I, D) as sorted int32 arrays, like ``utils._init_bc`` (utils.py:295-320)."""
    if (I is None) == (D is None):
        raise Exception("Give only I or only D!" if I is not None
                        else "Either I or D must be given!")
    allr = np.arange(n, dtype=np.int32)
    if D is None:
        I = np.unique(np.asarray(I, dtype=np.int32))
        D = np.setdiff1d(allr, I)
    else:
        D = np.unique(np.asarray(D, dtype=np.int32))
        I = np.setdiff1d(allr, D)
    return I.astype(np.int32), D.astype(np.int32)


def condense_device(A, b=None, x=None, I=None, D=None, expand=True):
    """``condense`` (``utils.py:540-603``) on a ``DeviceCSR``: returns
    ``A[I][:, I]`` (a canonical ``DeviceCSR``), ``b[I] - A[I][:, D] @ x[D]``
    and, with ``expand``, the full ``x`` and ``I`` (torch tensors on the
    device) for ``solve_device``."""
    import ctypes
    import torch
    from . import _lib
    from .device import DeviceCSR
    lib = _lib.load()
    dev = A.data.device
    n = A.shape[0]
    I_np, D_np = _dofs_split(n, I, D)
    colmap_np = np.full(n, -1, dtype=np.int32)
    colmap_np[I_np] = np.arange(len(I_np), dtype=np.int32)
    rows = torch.as_tensor(I_np, device=dev)
    colmap = torch.as_tensor(colmap_np, device=dev)
    nsel = len(I_np)
    out_indptr = torch.empty(nsel + 1, dtype=torch.int32, device=dev)
    wsb = ctypes.c_size_t(0)
    _lib.check(lib.femhip_csr_select_workspace(nsel, ctypes.byref(wsb)))
    ws = torch.empty(max(int(wsb.value), 1), dtype=torch.uint8, device=dev)
    nnz = ctypes.c_int64(0)
    _lib.check(lib.femhip_csr_select_pattern(
        _lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(rows), nsel, _lib.ptr(colmap),
        _lib.ptr(out_indptr), ctypes.byref(nnz), _lib.ptr(ws), int(wsb.value),
        _lib.stream_ptr()))
    out_indices = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)[:nnz.value]
    out_data = torch.empty(max(nnz.value, 1), dtype=torch.float64, device=dev)[:nnz.value]
    xt = None
    if x is not None:
        xt = torch.as_tensor(x, dtype=torch.float64, device=dev)
    bt = bout = None
    if b is not None:
        bt = torch.as_tensor(b, dtype=torch.float64, device=dev)
        bout = torch.empty(nsel, dtype=torch.float64, device=dev)
    _lib.check(lib.femhip_csr_select_fill(
        _lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(A.data), _lib.ptr(rows), nsel,
        _lib.ptr(colmap), _lib.ptr(out_indptr), _lib.ptr(out_indices), _lib.ptr(out_data),
        _lib.ptr(bt), _lib.ptr(xt), _lib.ptr(bout), _lib.stream_ptr()))
    AII = DeviceCSR(out_indptr, out_indices, out_data, (nsel, nsel))
    ret = (AII,) if b is None else (AII, bout)
    if expand:
        if xt is None:
            xt = torch.zeros(n, dtype=torch.float64, device=dev)
        ret += (xt, rows)
    return ret if len(ret) > 1 else ret[0]


def spmv(A, x):
    """``A @ x`` for a square ``DeviceCSR`` and a device vector."""
    import torch
    from . import _lib
    lib = _lib.load()
    y = torch.empty(A.shape[0], dtype=torch.float64, device=A.data.device)
    _lib.check(lib.femhip_spmv(_lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(A.data),
                               A.shape[0], _lib.ptr(x), _lib.ptr(y), _lib.stream_ptr()))
    return y


def solve_device(A, b, x=None, I=None, rtol=1e-10, atol=0.0, maxiter=None, x0=None,
                 check_every=16, info=None):
    """Solve ``A u = b`` on the device by Jacobi-preconditioned conjugate
    gradients (the algorithm of ``solver_iter_pcg``, ``utils.py:121-169``,
    with a tighter default tolerance).  With ``x`` and ``I`` (as returned by
    ``condense_device``) the solution is expanded into a copy of ``x``, like
    ``solve_linear`` (``utils.py:210-238``).  ``info`` (a dict) receives the
    iteration count and the final residual norm."""
    import ctypes
    import torch
    from . import _lib
    lib = _lib.load()
    dev = A.data.device
    n = A.shape[0]
    bt = torch.as_tensor(b, dtype=torch.float64, device=dev).contiguous()
    u = (torch.zeros(n, dtype=torch.float64, device=dev) if x0 is None
         else torch.as_tensor(x0, dtype=torch.float64, device=dev).clone())
    wsb = int(lib.femhip_pcg_workspace_size(n))
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    iters = ctypes.c_int64(0)
    resid = ctypes.c_double(0.0)
    _lib.check(lib.femhip_pcg(
        _lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(A.data), n, _lib.ptr(bt),
        _lib.ptr(u), float(rtol), float(atol), int(maxiter if maxiter is not None else 10 * n),
        int(check_every), _lib.ptr(ws), wsb, ctypes.byref(iters), ctypes.byref(resid),
        _lib.stream_ptr()))
    if info is not None:
        info['iterations'] = int(iters.value)
        info['residual'] = float(resid.value)
    if x is not None and I is not None:
        y = torch.as_tensor(x, dtype=torch.float64, device=dev).clone()
        y[torch.as_tensor(I, device=dev).long()] = u
        return y
    return u
