"""Recognition of user-written forms.

A user form written with ``femhip.helpers`` (``dot``, ``grad``, ...) is
called once with symbolic stand-ins for ``u``, ``v`` and ``w``; the
resulting expression is matched against the built-in kernels:

* ``dot(grad(u), grad(v))``                 -> laplace
* ``u * v``                                 -> mass
* ``dot(w.k * grad(u), grad(v))``           -> kappa_laplace (coefficient k)
* ``ddot(grad(u), grad(v))``                -> vector_laplace
* ``w.c * u * v`` (factors in any order)    -> coef_mass (coefficient c)
* ``v``                                     -> unit_load
* ``w.c * v``                               -> coef_load (coefficient c)

each possibly times numeric constants (the result is scaled; constant
factors do not change which entries are exactly zero).  Built-in models
carry their kernel name directly and are not traced.
"""
import numbers


class Expr:
    def __mul__(self, other):
        return Mul(self, wrap(other))

    def __rmul__(self, other):
        return Mul(wrap(other), self)

    def __add__(self, other):
        raise NotImplementedError("sums of terms are not accelerated")

    __radd__ = __sub__ = __rsub__ = __add__

    def __neg__(self):
        return Mul(Const(-1.0), self)

    def __truediv__(self, other):
        if isinstance(other, numbers.Number):
            return Mul(self, Const(1.0 / other))
        raise NotImplementedError("division by a field is not accelerated")


class Const(Expr):
    def __init__(self, c):
        self.c = float(c)


class Field(Expr):
    """Trial (``u``) or test (``v``) function value; ``.grad`` its gradient."""

    def __init__(self, which):
        self.which = which

    @property
    def grad(self):
        return Grad(self)

    @property
    def value(self):
        return self


class Grad(Expr):
    def __init__(self, f):
        self.f = f


class Coef(Expr):
    def __init__(self, name):
        self.name = name

    @property
    def value(self):
        return self


class Mul(Expr):
    def __init__(self, a, b):
        self.a, self.b = a, b


class Dot(Expr):
    def __init__(self, a, b):
        self.a, self.b = a, b


class DDot(Expr):
    def __init__(self, a, b):
        self.a, self.b = a, b


class W(dict):
    """Stand-in for the form's ``w`` (attribute and item access)."""

    def __getitem__(self, k):
        return Coef(k)

    def __getattr__(self, k):
        if k.startswith('__'):
            raise AttributeError(k)
        return Coef(k)

    def get(self, k, default=None):
        return Coef(k)


def wrap(x):
    if isinstance(x, Expr):
        return x
    if isinstance(x, numbers.Number):
        return Const(x)
    raise NotImplementedError("unsupported operand %r in a traced form" % (type(x),))


def _factors(e, out):
    if isinstance(e, Mul):
        _factors(e.a, out)
        _factors(e.b, out)
    else:
        out.append(e)
    return out


def _split(e):
    """(constant, list of non-constant factors) of a product."""
    c, rest = 1.0, []
    for f in _factors(e, []):
        if isinstance(f, Const):
            c *= f.c
        else:
            rest.append(f)
    return c, rest


def identify(form, nargs):
    """(kernel name, constant, params) of ``form`` or NotImplementedError."""
    if form is None:
        raise ValueError("form is not defined")
    spec = getattr(form, '_spec', None)
    if spec is not None:
        return spec
    u, v, w = Field('u'), Field('v'), W()
    try:
        expr = form(u, v, w) if nargs == 3 else form(v, w)
    except NotImplementedError:
        raise
    except Exception as exc:
        raise NotImplementedError(
            "form %r is not one of the accelerated kernels (%s)"
            % (getattr(form, '__name__', form), exc))
    expr = wrap(expr)
    c, fs = _split(expr)
    fields = [f for f in fs if isinstance(f, Field)]
    coefs = [f for f in fs if isinstance(f, Coef)]
    if nargs == 2:
        if len(fs) == 1 and isinstance(fs[0], Field) and fs[0].which == 'v':
            return ('unit_load', c, {})
        if (len(fs) == 2 and len(fields) == 1 and fields[0].which == 'v'
                and len(coefs) == 1):
            return ('coef_load', c, {'coef': coefs[0].name})
    else:
        if (len(fs) == 2 and all(isinstance(f, Field) for f in fs)
                and {f.which for f in fs} == {'u', 'v'}):
            return ('mass', c, {})
        if (len(fs) == 3 and len(fields) == 2 and len(coefs) == 1
                and sorted(f.which for f in fields) == ['u', 'v']):
            return ('coef_mass', c, {'coef': coefs[0].name})
        if (len(fs) == 1 and isinstance(fs[0], DDot) and isinstance(fs[0].a, Grad)
                and isinstance(fs[0].b, Grad) and {fs[0].a.f.which, fs[0].b.f.which} == {'u', 'v'}):
            return ('vector_laplace', c, {})
        if len(fs) == 1 and isinstance(fs[0], Dot):
            ca, fa = _split(fs[0].a)
            cb, fb = _split(fs[0].b)
            c *= ca * cb
            ga = [f for f in fa if isinstance(f, Grad)]
            gb = [f for f in fb if isinstance(f, Grad)]
            ka = [f for f in fa if isinstance(f, Coef)]
            if (len(ga) == 1 and len(gb) == 1 and len(fb) == 1
                    and {ga[0].f.which, gb[0].f.which} == {'u', 'v'}):
                if len(fa) == 1:
                    return ('laplace', c, {})
                if len(fa) == 2 and len(ka) == 1 and ga[0].f.which == 'u':
                    return ('kappa_laplace', c, {'coef': ka[0].name})
    raise NotImplementedError("form %r is not one of the accelerated kernels"
                              % getattr(form, '__name__', form))
