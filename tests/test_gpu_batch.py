"""Batched solver: every instance of a batch matches a plain solve of the
same instance (status, objective, iterate) and the oracle."""
import numpy as np
import pytest

import qp_libs as Q
from test_gpu_parity import build

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    return hprqp_amd


def test_batch_matches_plain_and_oracle(H):
    qps = Q.batched_instances(count=24, seed=11, n_lo=20, n_hi=400)
    models = [build(H, qp) for qp in qps]
    print("batch of", len(models), flush=True)
    res, secs = H.solve_batch(models, H.Parameters(verbose=False, stoptol=1e-6), bodies_per_launch=50)
    for k, (qp, model, r) in enumerate(zip(qps, models, res)):
        plain = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-6))
        o = Q.OracleSession(qp, stoptol=1e-6)
        orr, x, y, z, w = o.solve()
        assert r.status == plain.status == "OPTIMAL" == Q.STATUS[orr.status], (k, r.status, plain.status)
        assert abs(r.iter - orr.iter) <= 10, (k, r.iter, orr.iter)
        assert abs(r.objective - orr.primal_obj) <= 1e-6 * max(1.0, abs(orr.primal_obj)), k
        assert np.max(np.abs(r.x - x)) <= 1e-5 * max(1.0, np.max(np.abs(x))), k


def test_batch_fixture_traces(H):
    """Reference fixtures (tests/test_support.c) as one batch at max_iter
    100: each instance's iterate equals the reference's golden trace."""
    gold = np.load(Q.ROOT / "tests" / "golden" / "reference_traces.npz")
    names = ["STANDARD", "SCALE", "RESIDUALS", "DENSE_Q", "CONTROL_NONDIAG", "TALL_DIAG", "GPU_ITERATION_STANDARD",
             "EMPTY_A"]
    names += ["STANDARD_NOC"]
    F = Q.reference_fixtures()
    models = [build(H, F[n]) for n in names]
    res, _ = H.solve_batch(models, H.Parameters(verbose=False, max_iter=100), bodies_per_launch=7)
    for n, r in zip(names, res):
        want = gold[f"{n}/100/scalars"]
        assert r.iter == int(want[1]), n
        assert Q.STATUS_CODE[r.status] == int(want[0]), (n, r.status)
        scale = max(1.0, float(np.max(np.abs(gold[f"{n}/100/x"]))))
        assert np.max(np.abs(r.x - gold[f"{n}/100/x"])) <= 1e-10 * scale, n
