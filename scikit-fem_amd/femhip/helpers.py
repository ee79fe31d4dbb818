"""Form helpers with the names of ``skfem.helpers`` (``helpers.py:71-150``)
that work on the symbolic stand-ins used to recognise user forms."""
from . import trace


def grad(u):
    return u.grad


def dot(u, v):
    return trace.Dot(trace.wrap(u), trace.wrap(v))


def ddot(u, v):
    return trace.DDot(trace.wrap(u), trace.wrap(v))


def sym_grad(u):
    raise NotImplementedError("use femhip.models.linear_elasticity for sym_grad forms")
