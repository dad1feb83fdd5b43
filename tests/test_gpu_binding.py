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
