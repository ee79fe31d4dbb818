"""ctypes binding of ``libfemhip.so`` (the C ABI declared in
``include/femhip.h``).

torch is imported first so that the HIP runtime torch ships is the one the
library binds to (both carry the SONAME ``libamdhip64.so.7``); device memory
and streams therefore come straight from torch.  There is no fallback: when
the library is missing every accelerated entry point raises.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime before the library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libfemhip.so')

OK = 0
EINVAL = -1
EUNSUPPORTED = -2
EOVERFLOW = -3
EHIP = -4
EDEGENERATE = -5
ESIZE = -6

TRI, TET, HEX = 0, 1, 2
LAPLACE, MASS, KAPPA_LAPLACE, ELASTICITY, UNIT_LOAD, COEF_MASS, COEF_LOAD = 0, 1, 2, 3, 4, 5, 6

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double


class Geom(ctypes.Structure):
    _fields_ = [('kind', _i32), ('dim', _i32), ('nverts', _i32), ('nq', _i32),
                ('nelems', _i64), ('nnodes', _i64), ('p', _p), ('t', _p),
                ('W', _p), ('gdphi', _p)]


class Space(ctypes.Structure):
    _fields_ = [('nbs', _i32), ('vector', _i32), ('nbfun', _i32),
                ('reserved', _i32), ('ndofs', _i64), ('element_dofs', _p),
                ('phi', _p), ('dphi', _p)]


class Form(ctypes.Structure):
    _fields_ = [('kind', _i32), ('kappa_mode', _i32), ('kappa', _p),
                ('lam', _f64), ('two_mu', _f64)]


class FacetGeom(ctypes.Structure):
    _fields_ = [('kind', _i32), ('dim', _i32), ('degree', _i32), ('nq', _i32),
                ('nfacets', _i64), ('nnodes', _i64), ('nelems', _i64), ('p', _p),
                ('facets', _p), ('t', _p), ('tind', _p), ('X', _p), ('W', _p)]


class FusedInfo(ctypes.Structure):
    _fields_ = [('nparts', _i32), ('max_recs', _i32), ('nrows', _i32), ('max_slots', _i32),
                ('max_nodes', _i32), ('max_rows', _i32), ('rec_words', _i32),
                ('idbits', _i32), ('offbits', _i32), ('block', _i32),
                ('pruned', _i32), ('nrowrecs', _i32), ('nrecs', _i64), ('nlocal', _i64),
                ('dup', _f64)]


class FusedDev(ctypes.Structure):
    _fields_ = [('parts', _p), ('nodemap', _p), ('rows', _p), ('recs', _p),
                ('rec_elem', _p), ('info', FusedInfo)]


class FusedState(ctypes.Structure):
    _fields_ = [('s', _i32), ('sv', _i32), ('valid', _i32), ('s_first', _i32),
                ('sv_first', _i32), ('reserved0', _i32), ('epoch', ctypes.c_uint64),
                ('reserved1', ctypes.c_uint32 * 24),
                ('maxk', ctypes.c_uint64), ('maxv', ctypes.c_uint64),
                ('flags', _i32), ('checked', _i32),
                ('nsusp', _i64), ('mismatch', _i64), ('zeros', _i64),
                ('reserved3', ctypes.c_uint32 * 20)]


class FusedCheck(ctypes.Structure):
    _fields_ = [('t', _p), ('dphi', _p), ('W', _p), ('nelems', _i64), ('Q', _i32),
                ('reserved', _i32), ('zero_mask', _p), ('nzeros', _i64), ('queue', _p),
                ('capacity', _i64), ('verdict', _p)]


FUSED_BAD = 1
FUSED_OVERFLOW = 2


class FemhipError(RuntimeError):
    """HIP runtime failure reported by the library."""


# Exported symbols and their signatures: (restype, argtypes).
SIGNATURES = {
    'femhip_last_error': (ctypes.c_char_p, []),
    'femhip_version': (ctypes.c_int, []),
    'femhip_geometry_size': (_i64, [ctypes.POINTER(Geom)]),
    'femhip_geometry': (ctypes.c_int, [ctypes.POINTER(Geom), _p, _p, _p]),
    'femhip_local_bilinear': (ctypes.c_int, [ctypes.POINTER(Geom),
                                             ctypes.POINTER(Space),
                                             ctypes.POINTER(Form), _p, _p, _p]),
    'femhip_local_bilinear_mixed': (ctypes.c_int, [ctypes.POINTER(Geom),
                                                   ctypes.POINTER(Space),
                                                   ctypes.POINTER(Space),
                                                   ctypes.POINTER(Form), _p, _p, _p]),
    'femhip_local_linear': (ctypes.c_int, [ctypes.POINTER(Geom),
                                           ctypes.POINTER(Space),
                                           ctypes.POINTER(Form), _p, _p, _p]),
    'femhip_pattern_workspace': (ctypes.c_int, [_i64, ctypes.POINTER(ctypes.c_size_t)]),
    'femhip_pattern_build': (ctypes.c_int, [_p, _p, _i32, _i32, _i64, _i64, _i64,
                                            _p, _p, _p, _p, _p,
                                            ctypes.POINTER(_i64), _p,
                                            ctypes.c_size_t, _p]),
    'femhip_rows_plan_build': (ctypes.c_int, [_p, _i32, _i64, _i64, _p, _p, _p,
                                              ctypes.c_size_t, _p]),
    'femhip_segment_sum': (ctypes.c_int, [_p, _p, _p, _i64, _p, _p]),
    'femhip_zero_mask': (ctypes.c_int, [_p, _i64, _p, _p]),
    'femhip_zero_mask_mismatch': (ctypes.c_int, [_p, _i64, _p, _p, _p]),
    'femhip_fused_plan_build': (ctypes.c_int, [_i32, _i32, _i64, _i64, _p, _p, _p, _p,
                                               _i32, _i32, _i32, _i32,
                                               ctypes.POINTER(ctypes.c_void_p),
                                               ctypes.POINTER(FusedInfo)]),
    'femhip_fused_plan_export': (ctypes.c_int, [_p, _p, _p, _p, _p, _p]),
    'femhip_fused_plan_free': (None, [_p]),
    'femhip_fused_lds_bytes': (ctypes.c_size_t, [ctypes.POINTER(FusedInfo), _i32, _i32]),
    'femhip_fused_assemble': (ctypes.c_int, [ctypes.POINTER(FusedDev), _p, _i32, _i32, _p,
                                             _i64, _p, _p, _p, ctypes.POINTER(FusedCheck),
                                             _i32, _p]),
    'femhip_fused_static_queue': (_i64, []),
    'femhip_fused_laplace_mass': (ctypes.c_int, [ctypes.POINTER(FusedDev), _p, _p, _i32, _p,
                                                 _i64, _p, _p, _p, ctypes.POINTER(FusedCheck),
                                                 _i32, _p]),
    'femhip_rows_fused_plan_build': (ctypes.c_int, [_i32, _i32, _i64, _i64, _p, _p, _p, _p,
                                                    _i32, _i32, _i32,
                                                    ctypes.POINTER(ctypes.c_void_p),
                                                    ctypes.POINTER(FusedInfo)]),
    'femhip_rows_fused_plan_export': (ctypes.c_int, [ctypes.c_void_p, _p]),
    'femhip_rows_fused_lds_bytes': (ctypes.c_size_t, [ctypes.POINTER(FusedInfo)]),
    'femhip_rows_metric': (ctypes.c_int, [ctypes.POINTER(Geom), _p, _p, _p]),
    'femhip_rows_fused_assemble': (ctypes.c_int, [ctypes.POINTER(FusedDev), _p, _p, _i32,
                                                  ctypes.POINTER(Geom), ctypes.POINTER(Space),
                                                  _p, _p, _p, ctypes.POINTER(FusedCheck), _p]),
    'femhip_fused_adjoint_lds_bytes': (ctypes.c_size_t, [ctypes.POINTER(FusedInfo), _i32]),
    'femhip_fused_adjoint': (ctypes.c_int, [ctypes.POINTER(FusedDev), _p, _i32, _p, _i64, _p,
                                            _p, _p, _p, _p, _p, _p, _p]),
    'femhip_fused_assemble_static': (ctypes.c_int, [ctypes.POINTER(FusedDev), _i32, _i32, _p,
                                                    _i64, _p, _p, _p,
                                                    ctypes.POINTER(FusedCheck), _i32, _p]),
    'femhip_batched_kappa_values': (ctypes.c_int, [_p, _p, _p, _i64, _i64, _p, _i64, _p,
                                                   _p]),
    'femhip_hex2_elasticity_mfma': (ctypes.c_int, [ctypes.POINTER(Geom), ctypes.POINTER(Space),
                                                   _f64, _f64, _p, _p, _i32, _p]),
    'femhip_segment_sum_em': (ctypes.c_int, [_p, _p, _p, _i64, _i64, _i64, _p, _p]),
    'femhip_row_sum_em_plan': (ctypes.c_int, [_p, _p, _p, _i64, _i64, _i64, _p, _p]),
    'femhip_row_sum_em': (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, _i64, _i32, _p, _p]),
    'femhip_enforce_rows': (ctypes.c_int, [_p, _p, _p, _p, _i64, _f64, _p, _p, _p, _p]),
    'femhip_facet_local': (ctypes.c_int, [ctypes.POINTER(FacetGeom), _i32, _p, _i32, _p, _p]),
    'femhip_facet_interpolate': (ctypes.c_int, [ctypes.POINTER(FacetGeom), _p, _p, _p, _p, _p]),
    'femhip_facet_global_coordinates': (ctypes.c_int, [ctypes.POINTER(FacetGeom), _p, _p, _p]),
    'femhip_global_coordinates': (ctypes.c_int, [ctypes.POINTER(Geom), _p, _p, _p, _p]),
    'femhip_csr_expand_components': (ctypes.c_int, [_p, _p, _p, _i64, _i64, _i32, _p, _p, _p, _p]),
    'femhip_mesh_parameters': (ctypes.c_int, [ctypes.POINTER(Geom), _p, _p, _p]),
    'femhip_facet_normals': (ctypes.c_int, [ctypes.POINTER(FacetGeom), _p, _p, _p, _p]),
    'femhip_interpolate': (ctypes.c_int, [ctypes.POINTER(Geom), ctypes.POINTER(Space), _p, _p,
                                          _p, _p, _p]),
    'femhip_csr_select_workspace': (ctypes.c_int, [_i64, ctypes.POINTER(ctypes.c_size_t)]),
    'femhip_csr_select_pattern': (ctypes.c_int, [_p, _p, _p, _i64, _p, _p,
                                                 ctypes.POINTER(_i64), _p, ctypes.c_size_t,
                                                 _p]),
    'femhip_csr_select_fill': (ctypes.c_int, [_p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p,
                                              _p, _p]),
    'femhip_spmv': (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _p]),
    'femhip_pcg_workspace_size': (ctypes.c_size_t, [_i64]),
    'femhip_pcg': (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _f64, _f64, _i64, _i32, _p,
                                  ctypes.c_size_t, ctypes.POINTER(_i64),
                                  ctypes.POINTER(_f64), _p]),
}

_lib = None


def load():
    """Load the library (once) and bind every symbol of ``SIGNATURES``."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libfemhip.so is not built (expected at %s); run "
            "`python scikit-fem_amd/csrc/build.py` or "
            "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    """Map a femhip status code to the reference's exception types."""
    if rc == OK:
        return
    msg = load().femhip_last_error().decode(errors='replace')
    if rc in (EINVAL, ESIZE):
        raise ValueError(msg)
    if rc == EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == EOVERFLOW:
        raise OverflowError(msg)
    if rc == EDEGENERATE:
        raise Exception("Zero Jacobian determinant")
    raise FemhipError(msg)


def stream_ptr():
    """The current torch stream as a hipStream_t."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device pointer of a torch tensor (None stays NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())
