"""The Python wrapper's end-to-end cases (the reference's
python/tests/test_hprqp.py:12-60) on the HIP path, and MPS models read by
build_from_mps solved on the GPU against the oracle on the same model."""
import math

import numpy as np
import pytest

import qp_libs as Q
from test_mps import GOLD, model_arrays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible: the HIP path must run on the GPU")
    return hprqp_amd


def test_read_from_mps_and_max_iter(H):
    model = H.read_from_mps(GOLD / "simple_qp.mps")
    result = H.optimize(model, H.Parameters(verbose=False, max_iter=0))
    assert result.status == "MAX_ITER"
    assert len(result.x) == 2 and len(result.y) == 2


def test_build_from_qabc_unconstrained(H):
    model = H.build_from_qabc([[2.0, 0.0], [0.0, 2.0]], [-4.0, -2.0], [], [], [], [-math.inf, -math.inf],
                              [math.inf, math.inf])
    result = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
    assert result.status == "OPTIMAL"
    assert result.objective == pytest.approx(-5.0, abs=1e-6)
    assert result.x[0] == pytest.approx(2.0, abs=1e-6) and result.x[1] == pytest.approx(1.0, abs=1e-6)
    assert len(result.y) == 0


@pytest.mark.parametrize("name", ["simple_qp.mps", "model.mps"])
def test_mps_models_solved_on_gpu_match_oracle(H, name):
    model = H.read_from_mps(GOLD / name)
    qp = model_arrays(model)
    print("mps", name, flush=True)
    r = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
    o, x, y, z, w = Q.OracleSession(qp, stoptol=1e-8).solve()
    assert r.status == Q.STATUS[o.status] == "OPTIMAL"
    assert r.objective == pytest.approx(o.primal_obj, rel=1e-7, abs=1e-9)
    np.testing.assert_allclose(r.x, x, rtol=1e-6, atol=1e-8)


def test_generated_mps_model_on_gpu(H, tmp_path):
    qp0 = Q.random_qp(seed=7, n=300, m_eq=60, m_in=90, a_row_nnz=6, b_rows=75, b_row_nnz=4)
    path = tmp_path / "gen.mps"
    Q.write_mps(qp0, path, seed=7)
    model = H.read_from_mps(path)
    qp = model_arrays(model)
    print("generated mps", flush=True)
    for k in (1, 50, 100):
        r = H.optimize(model, H.Parameters(verbose=False, max_iter=k))
        o, x, y, z, w = Q.OracleSession(qp, max_iter=k).solve()
        assert r.iter == o.iter and r.status == Q.STATUS[o.status]
        scale = max(1.0, float(np.max(np.abs(x))))
        assert np.max(np.abs(r.x - x)) <= 1e-10 * scale, k


def test_quadratic_constraint_mps_solved_on_gpu(H, tmp_path):
    """tests/test_mps_qcqp.cpp:104-117: min -x + 3 with .5x^2 + x <= 4,
    x in [0, 4], y = 0, read through QCMATRIX and solved as a row cone:
    x = 2, objective 1."""
    from test_qcqp import fixture
    path = tmp_path / "qc.mps"
    path.write_text(fixture("L", "1", "4", " X X .5\n"))
    model = H.read_from_mps(path)
    print("qcqp fixture", flush=True)
    r = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8, time_limit=30))
    assert r.status == "OPTIMAL" and r.residual <= 1e-8
    assert abs(r.x[0] - 2) < 1e-6 and abs(r.objective - 1) < 1e-6
    assert .5 * r.x[0] ** 2 + r.x[0] <= 4 + 1e-6


@pytest.mark.parametrize("seed", [11, 12])
def test_generated_quadratic_rows_on_gpu(H, tmp_path, seed):
    """A generated QCMATRIX model (lifted to row cones) on the GPU against the
    oracle on the same model, per iterate."""
    from test_qcqp import qc_mps
    path = tmp_path / "gqc.mps"
    qc_mps(path, seed=seed, n=60)
    model = H.read_from_mps(path)
    qp = model_arrays(model)
    assert qp.num_row_cones > 0
    print("generated qcqp", seed, flush=True)
    for k in (1, 50, 100):
        r = H.optimize(model, H.Parameters(verbose=False, max_iter=k))
        o, x, y, z, w = Q.OracleSession(qp, max_iter=k).solve()
        assert r.iter == o.iter and r.status == Q.STATUS[o.status]
        for got, want in ((r.x, x), (r.y, y), (r.z, z)):
            scale = max(1.0, float(np.max(np.abs(want))))
            assert np.max(np.abs(got - want)) <= 1e-10 * scale, k
    r = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
    o, x, y, z, w = Q.OracleSession(qp, stoptol=1e-8).solve()
    assert r.status == Q.STATUS[o.status] == "OPTIMAL"
    assert r.objective == pytest.approx(o.primal_obj, rel=1e-7, abs=1e-9)
    np.testing.assert_allclose(r.x, x, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("params,status", [({"stoptol": 1.0}, "OPTIMAL"), ({"time_limit": 0.0}, "TIME_LIMIT"),
                                           ({"max_iter": 0}, "MAX_ITER")])
def test_termination_at_iteration_zero(H, params, status):
    """The loop's stopping rules at iteration 0 (solver_backend.c: a KKT below
    stoptol = 1, a time limit of zero, max_iter = 0), as the oracle stops."""
    qp = Q.reference_fixtures()["STANDARD"]
    model = H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c, (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val),
                              qp.AL, qp.AU, qp.l, qp.u)
    print("termination", params, flush=True)
    r = H.optimize(model, H.Parameters(verbose=False, **params))
    o, x, y, z, w = Q.OracleSession(qp, **params).solve()
    assert r.status == Q.STATUS[o.status] == status
    if "time_limit" in params:
        # elapsed > 0 is first true when the microsecond clock ticks: at
        # iteration 0 or 1 on the GPU (an iteration takes microseconds), after
        # a few iterations of the CPU oracle on this three-variable model
        assert r.iter <= 1 and o.iter <= 10
        return
    assert r.iter == o.iter == 0
    np.testing.assert_allclose(r.x, x, rtol=1e-12, atol=1e-14)


def test_batch_time_limit_stops_after_a_chunk(H):
    """A batch checks its time limit between chunks: with a limit of zero it
    stops after the first chunk, instances not yet converged report
    TIME_LIMIT with the iterations they ran, converged ones their result."""
    models = [H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                                (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u)
              for qp in (Q.random_qp(seed=s, n=300, m_eq=60, m_in=90, a_row_nnz=6, b_rows=75, b_row_nnz=4)
                         for s in range(3))]
    b = H.Batch(models, H.Parameters(verbose=False, stoptol=1e-10, time_limit=0.0))
    print("batch time limit", flush=True)
    left = b.solve(bodies_per_launch=20)
    res = [b.result(k) for k in range(len(models))]
    assert left == sum(r.status == "TIME_LIMIT" for r in res) > 0
    for r in res:
        assert r.status in ("TIME_LIMIT", "OPTIMAL")
        if r.status == "TIME_LIMIT":
            assert 0 < r.iter <= 20
