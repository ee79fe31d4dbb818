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
