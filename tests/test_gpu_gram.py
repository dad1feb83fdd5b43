"""Equality-row (Gram) metric on the GPU (libhprqp_hip.so through its C ABI)
against the oracle's restatement and the reference's own Gram contract
(tests/polishing/test_gpu.py:14-66, fixtures in tests/polishing/fixtures.py):

* activation: taken for 'affine' and 'mixed' under REQUEST, declined for
  'dependent', 'no_equalities' and 'box' with the solve then bit for bit the
  metric-off solve; AUTO never takes a plain QP and takes the grid-control
  model; results.gram_active reports it;
* the trace audit: each device step equals a dense solve with G = E E' of
  the device's scaled A (1e-10, the contract's tolerance), dy on the metric
  rows is -step / (2 sigma) and y_bar elsewhere (s - raw) / (2 sigma beta);
* per-iterate agreement with the oracle (1e-10 relative over the first 100
  loop bodies; the device solves with W = L^-1 where the oracle substitutes
  with L, so the two differ by rounding), and full solves by status,
  objective (1e-7) and iteration count (within one check interval).
"""
import numpy as np
import pytest

import qp_libs as Q
from test_gpu_parity import build, close

pytestmark = pytest.mark.gpu

KINDS = ["affine", "mixed", "dependent", "no_equalities", "box"]


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible: the HIP path must run on the GPU")
    return hprqp_amd


def _opt(H, qp, **kw):
    return H.optimize(build(H, qp), H.Parameters(verbose=False, **kw))


@pytest.mark.parametrize("kind", KINDS)
def test_gram_activation_contract(H, kind):
    print(f"case {kind}", flush=True)
    qp, optimum = Q.gram_fixture_qp(kind)
    off = _opt(H, qp, gram_mode=0)
    on = _opt(H, qp, gram_mode=2)
    active = kind in ("affine", "mixed")
    assert not off.gram_active and on.gram_active == active
    for r in (off, on):
        assert r.status == "OPTIMAL" and 0 <= r.residual <= 1e-8
        assert abs(r.objective - optimum) < 1e-7
    if not active:
        assert on.iter == off.iter
        for v in "xyzw":
            assert np.array_equal(getattr(on, v), getattr(off, v)), v
    if kind == "mixed":
        auto = _opt(H, qp)  # default gram_mode = AUTO
        assert not auto.gram_active
        for v in "xyzw":
            assert np.array_equal(getattr(auto, v), getattr(off, v)), v


def test_gram_mode_out_of_range(H):
    qp, _ = Q.gram_fixture_qp("affine")
    r = _opt(H, qp, gram_mode=3)
    assert r.status == "INVALID_ARGUMENT"


def _dense(ptr, ind, val, nrow, ncol):
    M = np.zeros((nrow, ncol))
    for r in range(nrow):
        lo, hi = int(ptr[r]), int(ptr[r + 1])
        np.add.at(M[r], ind[lo:hi].astype(np.int64), val[lo:hi])
    return M


def _audit(s, eq, rest, bodies):
    """After each advance, the last step's metric quantities against dense
    linear algebra on the device's scaled A."""
    n, m = s.model.num_variables(), s.model.num_constraints()
    A = _dense(*s.matrix("A"), m, n)
    G = A[eq] @ A[eq].T
    beta = s.scalar("gram_beta")
    lam = s.scalar("lambda_max_A")
    assert lam == 2 * beta
    worst = 0.0
    for k in bodies:
        s.advance(k)
        sigma = s.scalar("sigma")
        r, step = s.read("gram_r"), s.read("gram_step")
        want = np.linalg.solve(G, r)
        worst = max(worst, np.max(np.abs(step - want)) / (1 + np.max(np.abs(want))))
        dy, yb = s.read("dy"), s.read("y_bar")
        e = -step / (2 * sigma)
        worst = max(worst, np.max(np.abs(dy[eq] - e)) / (1 + np.max(np.abs(e))))
        if len(rest):
            er = (s.read("s")[rest] - s.read("Rp")[rest]) / (sigma * lam)
            worst = max(worst, np.max(np.abs(yb[rest] - er)) / (1 + np.max(np.abs(er))))
    return worst


@pytest.mark.parametrize("kind", ["affine", "mixed"])
def test_gram_trace_audit_fixtures(H, kind):
    qp, _ = Q.gram_fixture_qp(kind)
    s = H.Session(build(H, qp), H.Parameters(verbose=False, gram_mode=2))
    assert s.scalar("gram_active") == 1
    eq = np.flatnonzero(np.isfinite(qp.AL) & (qp.AL == qp.AU))
    rest = np.setdiff1d(np.arange(qp.m), eq)
    assert _audit(s, eq, rest, [1, 1, 5, 40]) < 1e-10


def test_gram_trace_audit_random(H):
    qp = Q.random_qp(seed=11, n=400, m_eq=150, m_in=100, a_row_nnz=8, b_rows=100, b_row_nnz=4)
    s = H.Session(build(H, qp), H.Parameters(verbose=False, gram_mode=2))
    assert s.scalar("gram_active") == 1 and s.scalar("gram_rows") == 150
    eq = np.arange(150)
    rest = np.arange(150, 250)
    assert _audit(s, eq, rest, [1, 3, 30, 100]) < 1e-10


@pytest.mark.parametrize("g", [3, 12])
def test_gram_trace_audit_grid(H, g):
    qp = Q.gram_fixture_spatial(g)
    s = H.Session(build(H, qp), H.Parameters(verbose=False))  # AUTO
    assert s.scalar("gram_active") == 2 and s.scalar("gram_rows") == g * g
    eq = np.arange(g * g)
    assert _audit(s, eq, np.arange(0), [1, 2, 50]) < 1e-10


@pytest.mark.parametrize("case", ["random", "grid8", "mixed"])
def test_gram_per_iterate_and_full_solve_match_oracle(H, case):
    print(f"case {case}", flush=True)
    if case == "random":
        qp, kw = Q.random_qp(seed=5, n=300, m_eq=100, m_in=80, a_row_nnz=8, b_rows=80, b_row_nnz=4), dict(gram_mode=2)
    elif case == "grid8":
        qp, kw = Q.gram_fixture_spatial(8), {}
    else:
        qp, kw = Q.gram_fixture_qp("mixed")[0], dict(gram_mode=2)
    s = H.Session(build(H, qp), H.Parameters(verbose=False, **kw))
    o = Q.OracleSession(qp, **kw)
    assert s.scalar("gram_active") == o.scalar("gram_active") > 0
    for name in ("sigma", "lambda_max_A", "gram_beta"):
        assert abs(s.scalar(name) - o.scalar(name)) <= 1e-12 * abs(o.scalar(name)), name
    for it in range(100):
        _, fin = o.advance(1)
        s.advance(1)
        if fin:
            break
        for v in ("x_bar", "y_bar", "w_bar", "x", "y", "w"):
            assert close(s.read(v), o.read(v)), (it, v)
    r = H.optimize(build(H, qp), H.Parameters(verbose=False, **kw))
    orr, x, y, z, w = Q.OracleSession(qp, **kw).solve()
    assert r.status == "OPTIMAL" and Q.STATUS[orr.status] == "OPTIMAL" and r.gram_active
    assert abs(r.objective - orr.primal_obj) <= 1e-7 * max(1.0, abs(orr.primal_obj))
    assert abs(int(r.iter) - int(orr.iter)) <= 100, (r.iter, orr.iter)


@pytest.mark.parametrize("g", [31, 63])
def test_gram_grid_auto_reaches_the_optimum(H, g):
    """The grid-control model at g = 31 (961 metric rows) and g = 63 (3,969,
    the size of the reference's live AUTO case, tests/polishing/test_gpu.py:68)
    solved with AUTO: the metric is taken and the solve ends OPTIMAL at the
    objective of the metric-off solve (1e-7; for g = 63 the metric-off solve
    runs to the same KKT tolerance)."""
    print(f"case g={g}", flush=True)
    qp = Q.gram_fixture_spatial(g)
    on = _opt(H, qp)
    assert on.gram_active and on.status == "OPTIMAL" and on.residual <= 1e-8
    off = _opt(H, qp, gram_mode=0, max_iter=200000)
    assert not off.gram_active and off.status == "OPTIMAL"
    assert abs(on.objective - off.objective) <= 1e-7 * (1 + abs(off.objective))
    assert on.iter < off.iter


def test_gram_closed_forms_unscaled_mixed_gpu(H):
    """'mixed' without scaling on the device: one metric row, G = [5], beta =
    (1 + 1e-12), sigma = sqrt(1.25^2 / 5 + 2 / beta) / ||c||_2, lambda_A =
    2 beta, and the first step's metric solve r / 5."""
    import math
    qp, _ = Q.gram_fixture_qp("mixed")
    s = H.Session(build(H, qp), H.Parameters(verbose=False, gram_mode=2, use_scaling=0))
    beta = 1 + 1e-12
    assert s.scalar("gram_active") == 1 and s.scalar("gram_rows") == 1
    assert s.scalar("gram_beta") == beta and s.scalar("lambda_max_A") == 2 * beta
    c = -np.array([[2.0, 0.3], [0.3, 1.0]]) @ [0.25, 0.5]
    want = math.sqrt(1.25 ** 2 / 5.0 + 2.0 / beta) / np.linalg.norm(c)
    assert abs(s.scalar("sigma") - want) <= 1e-15 * want
    s.advance(2)
    r, step = s.read("gram_r"), s.read("gram_step")
    assert r.shape == step.shape == (1,)
    assert abs(step[0] - r[0] / 5.0) <= 1e-15 * (1 + abs(r[0]))
