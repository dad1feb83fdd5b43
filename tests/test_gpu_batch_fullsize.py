"""The batched workload at its full size: 4,096 independent QPs (n ~
U{50..2000}, m = round(n/2) with 60 % equality rows, A at min(10, n) nnz per
row, Q = B'DB + 1e-2 I with B (n/4) x n at 4 nnz per row, box [-10, 10]) in
one batch at stoptol 1e-6, against the reference CPU path solving each
instance on its own (tests/golden/batched_4096.npz, written by
tests/golden/make_fullsize_golden.py from the reference built from its
sources):

* every instance ends OPTIMAL, as in the reference;
* every objective within 1e-6 relative (max(1, |objective|)) of the
  reference's, and every KKT residual below the stoptol;
* a sample of 16 instances solved one at a time by the plain GPU path gives
  the batch's status and objective.
"""
import hashlib

import numpy as np
import pytest

import fullsize as F
import qp_libs as Q
from test_gpu_parity import build

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import hprqp_amd as H
    if not H.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    gold = np.load(F.GOLD / "batched_4096.npz", allow_pickle=False)
    qps = Q.batched_instances(count=4096, seed=0)
    sha = hashlib.sha256()
    for qp in qps:
        sha.update(F.qp_sha256(qp).encode())
    assert sha.hexdigest() == str(gold["sha256"]), "generator output differs from the fixture's instances"
    models = [build(H, qp) for qp in qps]
    return H, qps, models, gold


def test_batched_4096_matches_reference(setup):
    H, qps, models, gold = setup
    print("batched_4096: creating", flush=True)
    b = H.Batch(models, H.Parameters(verbose=False, stoptol=1e-6))
    left = b.solve(100)
    info = b.status()
    print(f"batched_4096: solved in {b.seconds:.3f}s, left={left}, wave_row={b.wave_row}", flush=True)
    assert left == 0
    ref_st = [Q.STATUS[int(s)] for s in gold["status"]]
    assert info["status"] == ref_st and set(ref_st) == {"OPTIMAL"}
    obj, ref = info["objective"], gold["objective"]
    rel = np.abs(obj - ref) / np.maximum(1.0, np.abs(ref))
    it = info["iter"].astype(np.int64) - gold["iter"]
    worst = np.argsort(rel)[-5:][::-1]
    print("worst objective differences:", [(int(k), float(rel[k]), int(info["iter"][k]), int(gold["iter"][k]))
                                           for k in worst], flush=True)
    print(f"iterations: same {int(np.sum(it == 0))} / 4096, batch - reference: min {it.min()} max {it.max()} "
          f"sum {int(info['iter'].sum())} vs {int(gold['iter'].sum())}", flush=True)
    assert float(rel.max()) <= 1e-6, (int(worst[0]), float(rel.max()))
    assert float(info["kkt"].max()) <= 1e-6
    # a sample against the plain (one QP per call) GPU path
    params = H.Parameters(verbose=False, stoptol=1e-6)
    for k in range(0, 4096, 256):
        r = H.optimize(models[k], params)
        assert r.status == info["status"][k] == "OPTIMAL", k
        assert abs(r.objective - obj[k]) <= 1e-6 * max(1.0, abs(obj[k])), k
