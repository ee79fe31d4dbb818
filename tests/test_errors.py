"""Failure handling forced on the CPU: the host-side validation that runs
before anything is uploaded (out-of-range DOF indices, connectivity outside
the node array, non-finite coordinates), and the C ABI's argument checks
that return before any device work (undersized workspace, int32 overflow,
NULL arguments) with their status codes and messages."""
import ctypes

import numpy as np
import pytest

import femhip
from femhip import _lib
from femhip.device import check_host_inputs


def _basis():
    x = np.linspace(0, 1, 5)
    return femhip.Basis(femhip.MeshTet.init_tensor(x, x, x), femhip.ElementTetP1())


def test_out_of_range_dof_rejected():
    b = _basis()
    dofs = b.element_dofs.copy()
    dofs[2, 7] = b.N            # one past the end
    with pytest.raises(ValueError, match="DOF index out of range"):
        check_host_inputs(b.mesh.p, b.mesh.t, dofs, b.N)
    dofs[2, 7] = -1
    with pytest.raises(ValueError, match="DOF index out of range"):
        check_host_inputs(b.mesh.p, b.mesh.t, dofs, b.N)
    # through the basis: the custom numbering is refused before any upload
    b.dofs.element_dofs = dofs
    with pytest.raises(ValueError, match="DOF index out of range"):
        b.device()


def test_connectivity_out_of_range_rejected():
    b = _basis()
    t = b.mesh.t.copy()
    t[0, 0] = b.mesh.p.shape[1]
    with pytest.raises(ValueError, match="connectivity index out of range"):
        check_host_inputs(b.mesh.p, t, b.element_dofs, b.N)


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_coordinates_rejected(bad):
    b = _basis()
    b.mesh.p[1, 3] = bad
    with pytest.raises(ValueError, match="not finite"):
        b.device()


def test_valid_inputs_pass():
    b = _basis()
    check_host_inputs(b.mesh.p, b.mesh.t, b.element_dofs, b.N)


def _fake(n=1):
    # non-NULL pointers that the checks below never dereference
    return ctypes.c_void_p(0x1000 * n)


def test_c_abi_undersized_workspace():
    lib = _lib.load()
    nnz = ctypes.c_int64(0)
    rc = lib.femhip_pattern_build(_fake(1), _fake(2), 4, 4, 10, 20, 20, None, _fake(3),
                                  _fake(4), _fake(5), _fake(6), ctypes.byref(nnz),
                                  _fake(7), 16, None)
    assert rc == _lib.ESIZE
    assert b"workspace too small" in lib.femhip_last_error()


def test_c_abi_int32_overflow():
    lib = _lib.load()
    nnz = ctypes.c_int64(0)
    rc = lib.femhip_pattern_build(_fake(1), _fake(2), 10, 10, 1 << 26, 20, 20, None,
                                  _fake(3), _fake(4), _fake(5), _fake(6), ctypes.byref(nnz),
                                  _fake(7), 1 << 40, None)
    assert rc == _lib.EOVERFLOW
    nb = ctypes.c_size_t(0)
    assert lib.femhip_pattern_workspace(1 << 31, ctypes.byref(nb)) == _lib.EOVERFLOW


def test_c_abi_null_arguments():
    lib = _lib.load()
    nnz = ctypes.c_int64(0)
    rc = lib.femhip_pattern_build(None, None, 4, 4, 10, 20, 20, None, _fake(3), None, None,
                                  None, ctypes.byref(nnz), _fake(7), 1 << 30, None)
    assert rc == _lib.EINVAL
    assert lib.femhip_zero_mask(None, 5, None, None) == _lib.EINVAL
    assert lib.femhip_segment_sum(None, None, None, 3, None, None) == _lib.EINVAL
    with pytest.raises(ValueError):
        _lib.check(_lib.EINVAL)


def test_fused_structs_match_the_header():
    """The ctypes mirrors of the fused path's structs have the sizes the C
    side asserts (femhip_fused_state_t: 256 bytes, the atomically updated
    fields on the second 128-byte line)."""
    assert ctypes.sizeof(_lib.FusedState) == 256
    assert _lib.FusedState.maxk.offset == 128
    assert ctypes.sizeof(_lib.FusedInfo) == 12 * 4 + 3 * 8
    assert ctypes.sizeof(_lib.FusedCheck) == 3 * 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8


def test_fused_assemble_argument_checks():
    lib = _lib.load()
    buf = (ctypes.c_uint8 * 512)()
    st = ctypes.c_void_p((ctypes.addressof(buf) + 127) // 128 * 128)   # 128-byte aligned
    assert lib.femhip_fused_assemble(None, st, 0, 3, _fake(1), 10, None, _fake(2), None,
                                     None, 0, None) == _lib.EINVAL
    info = _lib.FusedInfo()
    info.block = 512
    dev = _lib.FusedDev(None, None, None, None, None, info)
    rc = lib.femhip_fused_assemble(ctypes.byref(dev), ctypes.c_void_p(st.value + 8), 0, 3,
                                   _fake(1), 10, None, _fake(2), None, None, 0, None)
    assert rc == _lib.EINVAL and b"aligned" in lib.femhip_last_error()
    info.block = 256
    dev = _lib.FusedDev(None, None, None, None, None, info)
    rc = lib.femhip_fused_assemble(ctypes.byref(dev), st, 0, 3, _fake(1), 10, None, _fake(2),
                                   None, None, 0, None)
    assert rc == _lib.EINVAL and b"block" in lib.femhip_last_error()
