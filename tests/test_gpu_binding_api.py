"""The library as the reference's callers use it: through the opaque-handle
API of include/hprqp_binding.h (src/shared/binding.c: hprqp_model_*,
hprqp_params_set_*, hprqp_optimize, hprqp_results_*), from a plain C client
compiled against include/ (tests/c_client/client.c) and through ctypes, the
way the reference's Python wrapper binds it (python/src/hprqp/__init__.py).
Expected values: the cases of python/tests/test_hprqp.py:12-60 (x = (2, 1),
objective -5; simple_qp.mps at max_iter 0), closed-form optima, and the
oracle on the same models."""
import subprocess

import numpy as np
import pytest

import qp_libs as Q

pytestmark = pytest.mark.gpu

CLIENT = Q.ROOT / "tests" / "c_client" / "_build" / "client"
SIMPLE = Q.ROOT / "tests" / "golden" / "mps" / "simple_qp.mps"


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible: the HIP path must run on the GPU")
    return hprqp_amd


@pytest.fixture(scope="module")
def B(H):
    return Q.binding_lib(Q.PKG / "libhprqp_hip.so")


def test_c_client_cases(H):
    assert CLIENT.exists(), "tests/c_client/_build/client is not built (run __graft_entry__.build())"
    print("c client", flush=True)
    out = subprocess.run([str(CLIENT), str(SIMPLE)], capture_output=True, text=True, timeout=100)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = {ln.split()[0]: dict(kv.split("=", 1) for kv in ln.split()[1:]) for ln in out.stdout.splitlines()}
    u = lines["binding_unconstrained"]
    assert u["rc"] == u["status"] == "OPTIMAL" and u["m"] == "0" and u["n"] == "2"
    assert abs(float(u["objective"]) + 5.0) <= 1e-6
    assert abs(float(u["x0"]) - 2.0) <= 1e-6 and abs(float(u["x1"]) - 1.0) <= 1e-6
    assert lines["binding_mps_read"] == {"rc": "OK", "n": "2", "m": "2"}
    mi = lines["binding_mps_max_iter"]
    assert mi["rc"] == mi["status"] == "MAX_ITER" and mi["n"] == "2" and mi["m"] == "2" and mi["iter"] == "0"
    # simple_qp.mps to optimality against the oracle on the model the MPS reader built
    opt = lines["binding_mps_optimal"]
    model = H.read_from_mps(SIMPLE)
    from test_mps import model_arrays
    o = Q.OracleSession(model_arrays(model), stoptol=1e-8).solve()[0]
    assert opt["status"] == Q.STATUS[o.status] == "OPTIMAL"
    assert abs(float(opt["objective"]) - o.primal_obj) <= 1e-7 * max(1.0, abs(o.primal_obj))
    assert float(opt["residual"]) <= 1e-8
    # python/tests/test_hprqp.py:45-60 through hprqp.h: x = (7/6, 5/6), objective 213/36
    e = lines["model_api_equality"]
    assert e["rc"] == e["status"] == "OPTIMAL"
    assert abs(float(e["objective"]) - 213 / 36) <= 1e-6
    assert abs(float(e["x0"]) - 7 / 6) <= 1e-6 and abs(float(e["x1"]) - 5 / 6) <= 1e-6
    # gram_mode through the C struct: REQUEST takes the metric on the equality
    # row and reaches the same optimum; AUTO leaves a plain QP alone
    ga, gr = lines["model_api_gram_auto"], lines["model_api_gram_request"]
    assert ga["status"] == gr["status"] == "OPTIMAL"
    assert ga["gram_active"] == "0" and gr["gram_active"] == "1"
    for g in (ga, gr):
        assert abs(float(g["objective"]) - 213 / 36) <= 1e-6
        assert abs(float(g["x0"]) - 7 / 6) <= 1e-6 and abs(float(g["x1"]) - 5 / 6) <= 1e-6
    assert lines["gpu_backend_available"]["value"] == "1"
    assert lines["model_api_cpu"]["rc"] == "NOT_IMPLEMENTED"


def test_binding_api_matches_model_api_and_oracle(H, B):
    """hprqp_optimize on a generated QP (arrays handed over as CSR views,
    binding.c:51-76) returns what solve() returns on the model built from
    the same arrays, bit for bit, and what the oracle returns to the parity
    bar: status, iterations, objective within 1e-7."""
    qp = Q.random_qp(seed=21, n=400, m_eq=120, m_in=80, a_row_nnz=8, heavy_rows=2, heavy_nnz=150, b_rows=100,
                     b_row_nnz=4)
    print("binding vs model api", flush=True)
    for kw in ({"stoptol": 1e-8}, {"max_iter": 37}, {"stoptol": 1e-6, "check_iter": 50, "use_scaling": 0}):
        rb = Q.binding_solve(B, qp, **kw)
        model = H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                                  (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u)
        rm = H.optimize(model, H.Parameters(verbose=False, **kw))
        assert rb["rc"] == rb["status"] == rm.status, kw
        assert rb["iter"] == rm.iter and rb["objective"] == rm.objective and rb["residual"] == rm.residual, kw
        for v in "xyzw":
            np.testing.assert_array_equal(rb[v], getattr(rm, v), err_msg=f"{v} {kw}")
        okw = {k: v for k, v in kw.items()}
        o = Q.OracleSession(qp, **okw)
        orr, x, y, z, w = o.solve()
        assert rb["status"] == Q.STATUS[orr.status], kw
        if "max_iter" in kw:
            assert rb["iter"] == orr.iter
            assert np.max(np.abs(rb["x"] - x)) <= 1e-10 * max(1.0, np.max(np.abs(x))), kw
        assert abs(rb["objective"] - orr.primal_obj) <= 1e-7 * max(1.0, abs(orr.primal_obj)), kw


def test_binding_rejects_invalid_models_like_the_reference(B):
    """A crossed bound and a NaN are rejected by hprqp_model_build_from_qabc_csr
    with INVALID_ARGUMENT (build.c:40-120), and hprqp_optimize on an unbuilt
    model returns INVALID_ARGUMENT without touching the device."""
    qp = Q.random_qp(seed=3, n=20, m_eq=4, m_in=4, a_row_nnz=3, b_rows=5, b_row_nnz=2)
    bad = Q.random_qp(seed=3, n=20, m_eq=4, m_in=4, a_row_nnz=3, b_rows=5, b_row_nnz=2)
    bad.l = bad.l.copy()
    bad.l[3] = bad.u[3] + 1.0
    assert Q.binding_solve(B, bad)["status"] == "INVALID_ARGUMENT"
    bad = Q.random_qp(seed=3, n=20, m_eq=4, m_in=4, a_row_nnz=3, b_rows=5, b_row_nnz=2)
    bad.c = bad.c.copy()
    bad.c[0] = np.nan
    assert Q.binding_solve(B, bad)["status"] == "INVALID_ARGUMENT"
    if Q.ref_available():
        for q in (qp,):
            assert Q.binding_solve(Q.ref_lib(), q, use_gpu=0)["status"] == "OPTIMAL"
    import ctypes as C
    vp = C.c_void_p
    model, prm, res = vp(), vp(), vp()
    B.hprqp_model_create(C.byref(model))
    B.hprqp_params_create(C.byref(prm))
    B.hprqp_results_create(C.byref(res))
    B.hprqp_params_set_verbose(prm, 0)
    try:
        assert B.hprqp_optimize(model, prm, res) == Q.STATUS_CODE["INVALID_ARGUMENT"]
    finally:
        B.hprqp_results_destroy(res)
        B.hprqp_params_destroy(prm)
        B.hprqp_model_destroy(model)
