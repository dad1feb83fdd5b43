"""Cone (SOC) path on the device against the reference's analytic SOC cases
and certificate (tests/gpu/test_gpu_soc.cu:152-273) and against the oracle's
restatement of the cone step, per iterate."""
import numpy as np
import pytest

import qp_libs as Q
from test_gpu_parity import build, close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    return hprqp_amd


@pytest.mark.parametrize("scaling", [0, 1])
@pytest.mark.parametrize("name", ["proj_cone", "affine_cone", "constant_cone", "constant_cone_infeasible"])
def test_soc_simple_models_gpu(H, name, scaling):
    qp, want = Q.soc_simple_models()[name]
    print(f"case {name} scaling={scaling}", flush=True)
    const = name.startswith("constant")
    r = H.optimize(build(H, qp), H.Parameters(verbose=False, use_scaling=scaling, print_frequency=0,
                                              stoptol=1e-8 if name == "constant_cone" else 1e-9,
                                              max_iter=1000 if const else 100000))
    if want is None:
        assert r.status == "MAX_ITER" and r.residual >= 0.49
        return
    assert r.status == "OPTIMAL", (r.status, r.iter, r.residual)
    assert np.max(np.abs(r.x - want)) <= 2e-7
    cert = Q.soc_certificate(qp, r.x, r.y, r.z)
    assert max(cert.values()) < 2e-6, cert


@pytest.mark.parametrize("q_kind", [0, 1, 2])
@pytest.mark.parametrize("scaling", [0, 1])
def test_soc_mixed_model_gpu(H, q_kind, scaling):
    qp = Q.soc_mixed_model(q_kind)
    print(f"mixed q={q_kind} scaling={scaling}", flush=True)
    r = H.optimize(build(H, qp), H.Parameters(verbose=False, use_scaling=scaling, stoptol=1e-9, print_frequency=0,
                                              max_iter=100000, time_limit=30.0))
    assert r.status == "OPTIMAL", (r.status, r.iter, r.residual)
    cert = Q.soc_certificate(qp, r.x, r.y, r.z)
    assert max(cert.values()) < 2e-6, cert
    if q_kind:
        assert np.max(np.abs(r.x - [0.25, -0.5, 5, 3, 4, 2, 0, 0])) <= 2e-5


def test_soc_per_iterate_matches_oracle(H):
    qp = Q.soc_qp(seed=1, n_box=300, cones=60, cone_dim=10, m=400, a_row_nnz=40)
    s = H.Session(build(H, qp), H.Parameters(verbose=False))
    o = Q.OracleSession(qp)
    assert abs(s.scalar("sigma") - o.scalar("sigma")) <= 1e-12 * o.scalar("sigma")
    for it in range(100):
        s.advance(1)
        o.advance(1)
        for v in ("x_bar", "y_bar", "w_bar", "z_bar"):
            assert close(s.read(v), o.read(v)), (it, v)


def test_soc_generated_full_solve(H):
    qp = Q.soc_qp(seed=2, n_box=500, cones=200, cone_dim=10, m=1000, a_row_nnz=40)
    r = H.optimize(build(H, qp), H.Parameters(verbose=False, stoptol=1e-8))
    o = Q.OracleSession(qp)
    orr, x, y, z, w = o.solve()
    assert r.status == "OPTIMAL" and Q.STATUS[orr.status] == "OPTIMAL"
    assert abs(r.objective - orr.primal_obj) <= 1e-7 * max(1.0, abs(orr.primal_obj))
    cert = Q.soc_certificate(qp, r.x, r.y, r.z)
    assert max(cert.values()) < 1e-6, cert


@pytest.mark.parametrize("case", ["var_cones", "mixed", "row_cones"])
def test_soc_fused_cone_kernel_matches_thread_per_cone(H, monkeypatch, case):
    """The cone step with L lanes per cone (projection and finish fused,
    coalesced) against one thread per cone followed by the finish map
    (HPRQP_CONE_LANES=0): the same bits after every loop body, for variable
    cones (dimension 10), the mixed model (variable and row cones of several
    sizes) and row cones of the grid-control model. The variable-cone and
    grid-control models have an exactly diagonal Q (the latter with empty
    rows too), so their default runs take the one-launch x step (Q w, the
    projections, the finish and Q w_bar elementwise) against the separate
    products of HPRQP_CONE_LANES=0; the mixed model's Q is not diagonal."""
    if case == "var_cones":
        qp = Q.soc_qp(seed=4, n_box=200, cones=40, cone_dim=10, m=300, a_row_nnz=20)
    elif case == "mixed":
        qp = Q.soc_mixed_model(2)
    else:
        qp = Q.gram_fixture_spatial(6)
    print(f"case {case}", flush=True)
    runs = {}
    for lanes in ("0", "auto"):
        if lanes == "0":
            monkeypatch.setenv("HPRQP_CONE_LANES", "0")
        else:
            monkeypatch.delenv("HPRQP_CONE_LANES", raising=False)
        s = H.Session(build(H, qp), H.Parameters(verbose=False, gram_mode=0))
        trace = []
        for _ in range(60):
            s.advance(1)
            trace.append({v: s.read(v) for v in ("x", "y", "w", "x_bar", "y_bar")})
        runs[lanes] = trace
    for k, (a, b) in enumerate(zip(runs["0"], runs["auto"])):
        for v in a:
            assert np.array_equal(a[v], b[v]), (k, v)
