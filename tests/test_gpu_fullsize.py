"""The plain path at the benchmark's full size (plain_n1M: n = 1e6, m = 5e5,
19.9M nnz in A with 5,000 rows of 2,000 nnz, 12.2M nnz in Q after
symmetrisation): setup, per-iterate and end-to-end parity with the reference
CPU path, plus size-independent properties.

The reference's full solve and the oracle's setup and sampled iterates are
committed in tests/golden/plain_n1M.json (tests/golden/make_fullsize_golden.py
ran the reference built from its sources, and the oracle, which matched it
bit for bit):

* lambda_max(AA') and lambda_max(Q) within 1e-10 relative of the oracle's
  (src/cpu/eigs.c:19-185), with the same power-iteration step counts, and
  sigma0;
* the device transpose equals a stable (column, row) sort of the device's
  scaled A, bit for bit; Q's and A's scaled values, the row / column norms
  and the scaled bounds equal the oracle's bit for bit;
* x_bar, y_bar, w_bar, x, y, w within 1e-10 relative of the oracle (run live
  from the fixture's lambdas) after loop bodies 1-10, 20, ..., 100, and of
  the fixture's sampled entries and norms at bodies 10 and 100;
* the full solve (stoptol 1e-8): the reference's status, objective within
  1e-7 relative, KKT below stoptol, and the iteration count within one check
  interval (100) of the reference's (src/shared/solver_backend.c:338-569);
* the column-panel carry leaves each row's sum sequential: the panelled
  products equal the unsplit ones bit for bit (HPRQP_PANELS forces counts).
"""
import os

import numpy as np
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
    fix = F.fixture("plain_n1M")
    qp = Q.random_qp(**Q.PLAIN_N1M)
    assert F.qp_sha256(qp) == fix["sha256"], "generator output differs from the fixture's instance"
    model = build(H, qp)
    return H, qp, model, fix


def test_fullsize_transpose_bit_exact(setup):
    H, qp, model, fix = setup
    s = H.Session(model, H.Parameters(verbose=False))
    rp, ci, v = s.matrix("A")
    trp, tci, tv = s.matrix("AT")
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.uint64), np.diff(rp.astype(np.int64)))
    order = np.lexsort((rows, ci))
    assert np.array_equal(trp[1:] - trp[:-1], np.bincount(ci.astype(np.int64), minlength=qp.n))
    assert np.array_equal(tci, rows[order])
    assert np.array_equal(tv, v[order])
    print("panels", s.scalar("panels_Q"), s.scalar("panels_A"), s.scalar("panels_AT"), flush=True)


def test_fullsize_setup_matches_oracle(setup):
    H, qp, model, fix = setup
    s = H.Session(model, H.Parameters(verbose=False))
    want = fix["oracle_setup"]
    print("plain_n1M setup", {k: s.scalar(k) for k in want}, flush=True)
    for k in ("power_steps_A", "power_steps_Q", "Q_is_diag", "noC"):
        assert s.scalar(k) == want[k], k
    for k in ("lambda_max_A", "lambda_max_Q"):
        assert abs(s.scalar(k) - want[k]) <= 1e-10 * want[k], (k, s.scalar(k), want[k])
    assert abs(s.scalar("sigma") - want["sigma"]) <= 1e-12 * want["sigma"]


def test_fullsize_first_100_bodies_match_oracle(setup):
    H, qp, model, fix = setup
    s = H.Session(model, H.Parameters(verbose=False))
    o = Q.OracleSession(qp, lambda_A_given=fix["oracle_setup"]["lambda_max_A"],
                        lambda_Q_given=fix["oracle_setup"]["lambda_max_Q"])
    assert s.scalar("sigma") == o.scalar("sigma")
    for name in ("col_norm", "row_norm", "c", "AL", "AU"):
        assert np.array_equal(s.read(name), o.read(name)), name
    for name in ("Q", "A"):
        assert np.array_equal(s.matrix(name)[2], o.matrix(name)[2]), name
    done = 0
    for body in list(range(1, 11)) + list(range(20, 101, 10)):
        s.advance(body - done)
        o.advance(body - done)
        done = body
        print(f"plain_n1M body {body}", flush=True)
        for v in ("x_bar", "y_bar", "w_bar", "x", "y", "w"):
            got, want = s.read(v), o.read(v)
            assert close(got, want), (body, v, F.rel_err(got, want))
            if str(body) in fix["oracle_iterates"] and v in ("x_bar", "y_bar", "w_bar"):
                F.check_sampled(got, fix["oracle_iterates"][str(body)][v], 1e-10)


def test_fullsize_solve_matches_reference(setup):
    H, qp, model, fix = setup
    ref = fix["reference"]
    r = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
    print(f"plain_n1M: gpu {r.status} iter={r.iter} iter_6={r.iter_6} obj={r.objective!r} kkt={r.residual:.3e} "
          f"solve={r.time:.3f}s total={r.total_time:.3f}s; reference {ref}", flush=True)
    assert r.status == ref["status"] == "OPTIMAL"
    assert abs(r.objective - ref["objective"]) <= 1e-7 * max(1.0, abs(ref["objective"]))
    assert r.residual <= 1e-8
    assert abs(int(r.iter) - int(ref["iter"])) <= 100


def test_fullsize_panel_products_bitwise(setup):
    H, qp, model, fix = setup
    outs = {}
    for P in ("1", "4"):
        os.environ["HPRQP_PANELS"] = P
        try:
            s = H.Session(model, H.Parameters(verbose=False, max_iter=3))
            s.advance(4)
            outs[P] = [s.read(v) for v in ("x_bar", "y_bar", "w_bar", "x", "w", "y")]
        finally:
            del os.environ["HPRQP_PANELS"]
    for a, b in zip(outs["1"], outs["4"]):
        assert close(a, b, 1e-12)


def test_fullsize_occupancy_cap_changes_no_bit(setup):
    """The residency cap (unused dynamic LDS per workgroup) runs the same
    blocks in the same order: iterates with it forced on and forced off are
    bitwise equal, at two column panels (the value the autotuner picks
    here) and without panels."""
    H, qp, model, fix = setup
    for P in ("2", "1"):
        outs = {}
        for cap in ("1", "0"):
            os.environ["HPRQP_PANELS"], os.environ["HPRQP_OCCUPANCY_CAP"] = P, cap
            try:
                s = H.Session(model, H.Parameters(verbose=False, max_iter=3))
                pads = [s.scalar("lds_pad_" + m) for m in ("Q", "A", "AT")]
                assert (min(pads) > 0) if cap == "1" else (max(pads) == 0), (P, cap, pads)
                s.advance(4)
                outs[cap] = [s.read(v) for v in ("x_bar", "y_bar", "w_bar", "x", "w", "y")]
                del s
            finally:
                del os.environ["HPRQP_PANELS"], os.environ["HPRQP_OCCUPANCY_CAP"]
        for a, b in zip(outs["1"], outs["0"]):
            assert np.array_equal(a, b), P
