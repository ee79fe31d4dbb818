"""Vector Q2 hexahedron linear elasticity on FP64 matrix cores: the second
assembly on the same inputs runs the MFMA local kernel; it must reproduce
the reference's pattern and values within 1e-12."""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import femhip
from femhip.models import linear_elasticity, lame_parameters

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
TOL = 1e-12
pytestmark = pytest.mark.gpu


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_mfma_vs_fixture():
    z = np.load(os.path.join(GOLD, 'small_vhex2_n4j.npz'))
    m = femhip.MeshHex(z['p'], z['t'])
    b = femhip.Basis(m, femhip.ElementVector(femhip.ElementHex2()), intorder=5)
    lam, mu = lame_parameters(1.0, 0.3)
    form = linear_elasticity(lam, mu)
    A1 = form.assemble_device(b)          # exact path
    A2 = form.assemble_device(b)          # MFMA path
    A3 = form.assemble_device(b)
    torch.cuda.synchronize()
    assert b.device().validated, "MFMA path not taken"
    d2, d3 = A2.data.cpu().numpy(), A3.data.cpu().numpy()
    assert np.array_equal(d2, d3)
    ref = z['linear_elasticity_data']
    assert np.array_equal(A2.indptr.cpu().numpy(), z['linear_elasticity_indptr'])
    assert np.array_equal(A2.indices.cpu().numpy(), z['linear_elasticity_indices'])
    err = np.abs(d2 - ref).max() / np.abs(ref).max()
    assert err <= TOL, err
    err1 = np.abs(A1.data.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err1 <= TOL, err1


def test_mfma_digest_11():
    dig = json.load(open(os.path.join(GOLD, 'digest_vhex2_n11j.json')))
    x = np.linspace(0, 1, 11)
    m = femhip.MeshHex.init_tensor(x, x, x).jittered(0.15, 0.1)
    assert _sha(m.p) == dig['p_sha']
    b = femhip.Basis(m, femhip.ElementVector(femhip.ElementHex2()), intorder=5)
    lam, mu = lame_parameters(1.0, 0.3)
    form = linear_elasticity(lam, mu)
    form.assemble_device(b)
    A = form.assemble_device(b)
    d = dig['linear_elasticity']
    assert A.nnz == d['nnz']
    assert _sha(A.indptr.cpu().numpy().astype(np.int32)) == d['indptr_sha']
    assert _sha(A.indices.cpu().numpy().astype(np.int32)) == d['indices_sha']
    data = A.data.cpu().numpy()
    slots = np.array(d['slots'])
    assert np.abs(data[slots] - np.array(d['values'])).max() <= TOL * d['max_abs']
