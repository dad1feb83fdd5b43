"""The cone workload at its full size (soc_n220k: 20,000 ten-dimensional
second-order cones after a 20,000-variable box prefix, 100,000 equality rows
at 40 nnz, diagonal Q, stoptol 1e-8) on the device against the oracle's
restatement of the cone step (src/gpu/soc_iteration.cuh:45-82; the
reference has no CPU cone path). The oracle's setup, sampled iterates and
full solve are committed in tests/golden/soc_n220k.json (written by
tests/golden/make_fullsize_golden.py); the per-iterate check also runs the
oracle live from the fixture's own lambda_max(AA'):

* lambda_max(AA') within 1e-10 relative, with the same power-iteration step
  count (2,586), and sigma0;
* x_bar, y_bar, w_bar, z_bar within 1e-10 relative after every one of the
  first 100 loop bodies, and against the fixture at bodies 10 and 100;
* the full solve: the oracle's status and iteration count, objective within
  1e-7 relative, KKT below stoptol, and the independent certificate of
  tests/gpu/test_gpu_soc.cu:152-196 below 1e-6.
"""
import pytest

import fullsize as F
import qp_libs as Q
from test_gpu_parity import build, close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import hprqp_amd as H
    if not H.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    fix = F.fixture("soc_n220k")
    qp = Q.soc_qp(**Q.SOC_N220K)
    assert F.qp_sha256(qp) == fix["sha256"], "generator output differs from the fixture's instance"
    return H, qp, build(H, qp), fix


def test_soc_fullsize_setup_matches_oracle(setup):
    H, qp, model, fix = setup
    s = H.Session(model, H.Parameters(verbose=False))
    want = fix["oracle_setup"]
    print("soc_n220k setup", {k: s.scalar(k) for k in want}, flush=True)
    assert s.scalar("power_steps_A") == want["power_steps_A"]
    assert abs(s.scalar("lambda_max_A") - want["lambda_max_A"]) <= 1e-10 * want["lambda_max_A"]
    assert s.scalar("lambda_max_Q") == want["lambda_max_Q"]  # diagonal Q: max |q_ii|
    assert s.scalar("Q_is_diag") == want["Q_is_diag"] == 1
    assert abs(s.scalar("sigma") - want["sigma"]) <= 1e-12 * want["sigma"]


def test_soc_fullsize_first_100_bodies_match_oracle(setup):
    H, qp, model, fix = setup
    print("soc_n220k per-iterate", flush=True)
    s = H.Session(model, H.Parameters(verbose=False))
    o = Q.OracleSession(qp, lambda_A_given=fix["oracle_setup"]["lambda_max_A"])
    assert o.scalar("lambda_max_A") == fix["oracle_setup"]["lambda_max_A"]
    for body in range(1, 101):
        s.advance(1)
        o.advance(1)
        for v in ("x_bar", "y_bar", "w_bar", "z_bar"):
            got = s.read(v)
            assert close(got, o.read(v)), (body, v, F.rel_err(got, o.read(v)))
            if str(body) in fix["oracle_iterates"] and v != "z_bar":
                F.check_sampled(got, fix["oracle_iterates"][str(body)][v], 1e-10)


def test_soc_fullsize_solve_matches_oracle_and_certificate(setup):
    H, qp, model, fix = setup
    want = fix["oracle_full"]
    r = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
    print(f"soc_n220k: gpu {r.status} iter={r.iter} obj={r.objective!r} kkt={r.residual:.3e} solve={r.time:.3f}s; "
          f"oracle {want}", flush=True)
    assert r.status == want["status"] == "OPTIMAL"
    assert r.iter == want["iter"]
    assert abs(r.objective - want["objective"]) <= 1e-7 * max(1.0, abs(want["objective"]))
    assert r.residual <= 1e-8
    cert = Q.soc_certificate(qp, r.x, r.y, r.z)
    print("certificate", cert, "oracle's", fix["oracle_certificate"], flush=True)
    assert max(cert.values()) < 1e-6, cert
