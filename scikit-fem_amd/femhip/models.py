"""Built-in forms (``skfem/models/poisson.py:7-24``,
``skfem/models/elasticity.py:7-53``) tagged with their device kernel."""
from .forms import BilinearForm, LinearForm


def _tagged(name, nargs, params=None, doc=""):
    def f(*args):
        raise NotImplementedError("built-in form %s runs on the device" % name)
    f.__name__ = name
    f.__doc__ = doc
    f._spec = (name, 1.0, params or {})
    return f


laplace = BilinearForm(_tagged('laplace', 3, doc="dot(grad(u), grad(v))"))
mass = BilinearForm(_tagged('mass', 3, doc="u * v"))
vector_laplace = BilinearForm(_tagged('vector_laplace', 3, doc="ddot(grad(u), grad(v))"))
unit_load = LinearForm(_tagged('unit_load', 2, doc="v"))
kappa_laplace = BilinearForm(_tagged('kappa_laplace', 3, {'coef': 'kappa'},
                                     doc="dot(w['kappa'] * grad(u), grad(v))"))


def lame_parameters(E, nu):
    """Lame parameters (lambda, mu) from Young's modulus and Poisson ratio."""
    return (E * nu / ((1. + nu) * (1. - 2. * nu)), E / (2. * (1. + nu)))


def linear_elasticity(Lambda=1., Mu=1.):
    """``ddot(C(sym_grad(u)), sym_grad(v))`` with
    ``C(T) = 2 Mu T + Lambda tr(T) I``."""
    return BilinearForm(_tagged('linear_elasticity', 3,
                                {'Lambda': float(Lambda), 'Mu': float(Mu)},
                                doc="linear elasticity"))
