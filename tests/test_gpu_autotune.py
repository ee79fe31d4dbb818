"""The fused kernel's autotuner: every candidate configuration computes the
same matrix, the fastest plan is installed on the basis and used by the
drop-in call, and the opt-in cache is keyed by the mesh and pattern (a
cached configuration is applied to that mesh only)."""
import json

import numpy as np
import pytest

import femhip
from femhip import autotune
from femhip.models import laplace
from oracle import skfem_np as O

pytestmark = pytest.mark.gpu


def test_tune_installs_best_plan_and_caches_per_mesh(tmp_path):
    x = np.linspace(0, 1, 21)
    m = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1 / 20)
    b = femhip.Basis(m, femhip.ElementTetP1())
    cache = str(tmp_path / 'tune.json')
    cands = [(4096, 768, 512, 1, 0), (2048, 400, 512, 0, 0), (4096, 768, 512, 1, 64),
             (6144, 1152, 1024, 1, 0)]
    print('tuning', flush=True)
    rec = autotune.tune_fused(b, 'laplace', candidates=cands, steps=3, cache=cache)
    assert len(rec['trials']) == len(cands)
    assert tuple(rec['best']) in cands
    d = b.device()
    fp = d.fused_plans[d.fused_last['laplace'].signature]
    assert fp.groups == rec['best'][4] and fp.caps == tuple(rec['best'][:2])
    A = laplace.assemble(b)                  # through the installed plan
    assert d.last_verdict[0] == 1
    R = O.bilinear_assemble(O.form_laplace, O.CellBasis(O.MeshTet(m.p, m.t), O.ElementTetP1()))
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    assert np.abs(A.data - R.data).max() <= 1e-12 * np.abs(R.data).max()
    db = json.load(open(cache))
    assert len(db) == 1
    key = next(iter(db))
    assert key == autotune.plan_key('tet', 'laplace', m.t, R.indptr, R.indices)
    # another mesh has another key: the entry is not applied to it
    m2 = femhip.MeshTet.init_tensor(x, x, x).jittered(0.2, 1 / 20, seed=4)
    R2 = O.bilinear_assemble(O.form_laplace, O.CellBasis(O.MeshTet(m2.p, m2.t), O.ElementTetP1()))
    assert autotune.lookup(autotune.plan_key('tet', 'laplace', m2.t, R2.indptr, R2.indices),
                           cache=cache) is None or np.array_equal(m2.t, m.t)
