"""femhip: MI355X-native assembly of scikit-fem's cell forms.

Drop-in names for the accelerated path of scikit-fem
(``BilinearForm.assemble`` / ``LinearForm.assemble`` on ``CellBasis``):

    from femhip import MeshTet, ElementTetP1, Basis
    from femhip.models import laplace, unit_load
    basis = Basis(MeshTet.init_tensor(x, x, x), ElementTetP1())
    A = laplace.assemble(basis)        # scipy.sparse.csr_matrix
    b = unit_load.assemble(basis)      # numpy.ndarray

The arithmetic runs in hand-written HIP kernels for gfx950 (libfemhip.so);
``assemble_device`` keeps the result in HBM as a ``DeviceCSR``.
"""
from .mesh import MeshTri, MeshTet, MeshHex, MeshTri1, MeshTet1, MeshHex1
from .element import (ElementTriP1, ElementTriP2, ElementTetP1, ElementTetP2,
                      ElementHex1, ElementHex2, ElementVector)
from .basis import CellBasis, Basis, FacetBasis, DiscreteField
from .forms import BilinearForm, LinearForm, asm
from .utils import enforce, enforce_device, condense_device, solve_device
from .device import DeviceCSR
from .mapping import MappingAffine, MappingIsoparametric

__version__ = "0.1.0"

__all__ = [
    'MeshTri', 'MeshTet', 'MeshHex', 'MeshTri1', 'MeshTet1', 'MeshHex1',
    'ElementTriP1', 'ElementTriP2', 'ElementTetP1', 'ElementTetP2',
    'ElementHex1', 'ElementHex2', 'ElementVector', 'CellBasis', 'Basis', 'FacetBasis',
    'BilinearForm', 'LinearForm', 'asm', 'enforce', 'DeviceCSR', 'MappingAffine',
    'MappingIsoparametric',
]
