"""Equality-row (Gram) metric of the oracle. The reference runs it on the GPU
only (src/gpu/qp_row_metric.cu, src/gpu/row_metric.cu, with a sparse direct
factor of G), so the oracle's restatement is pinned by what the reference's
own contract test checks (tests/polishing/test_gpu.py:14-66) and by what the
mathematics fixes:

* activation: the metric is taken for 'affine' and 'mixed', declined for
  'dependent' (singular G), 'no_equalities' and 'box', and a declined metric
  leaves the solve bit for bit as with the metric off; AUTO never takes a
  plain QP and takes the grid-control model;
* the trace audit of test_gpu.py:20-36: each metric step equals a dense
  solve with G, y_bar follows y - step / (2 sigma) on the metric rows and
  (s - raw) / (2 sigma beta) on the others, the quadratic equals
  2 dy_E' G dy_E + 2 beta ||dy_rest||^2; and G = E E';
* closed forms of beta, sigma and lambda_A on an unscaled model;
* optimality: every active solve reaches the analytic optimum (1e-7) or the
  objective of the same model solved without the metric.
"""
import math

import numpy as np
import pytest

import qp_libs as Q

KINDS = ["affine", "mixed", "dependent", "no_equalities", "box"]


def _solve(qp, **kw):
    o = Q.OracleSession(qp, print_frequency=0, **kw)
    assert o.status == 0
    r, x, y, z, w = o.solve()
    return o, r, (x, y, z, w)


@pytest.mark.parametrize("kind", KINDS)
def test_gram_activation_matches_reference_contract(kind):
    qp, optimum = Q.gram_fixture_qp(kind)
    o_off, r_off, sol_off = _solve(qp, gram_mode=0)
    o_on, r_on, sol_on = _solve(qp, gram_mode=2)
    active = kind in ("affine", "mixed")
    assert o_off.scalar("gram_active") == 0
    assert o_on.scalar("gram_active") == (1 if active else 0)
    for o, r in ((o_off, r_off), (o_on, r_on)):
        assert Q.STATUS[r.status] == "OPTIMAL" and 0 <= r.residuals <= 1e-8
        assert abs(r.primal_obj - optimum) < 1e-7
    if not active:
        assert r_on.iter == r_off.iter
        for a, b in zip(sol_on, sol_off):
            assert np.array_equal(a, b)
    if kind == "mixed":
        o_auto, r_auto, sol_auto = _solve(qp, gram_mode=1)
        assert o_auto.scalar("gram_active") == 0 and o_auto.scalar("gram_mode") == 0
        for a, b in zip(sol_auto, sol_off):
            assert np.array_equal(a, b)


def test_gram_mode_out_of_range_is_rejected():
    qp, _ = Q.gram_fixture_qp("affine")
    assert Q.OracleSession(qp, gram_mode=3).status == Q.STATUS_CODE["INVALID_ARGUMENT"]
    assert Q.OracleSession(qp, gram_mode=-1).status == Q.STATUS_CODE["INVALID_ARGUMENT"]


def test_gram_closed_forms_unscaled_mixed():
    """'mixed' without scaling: one metric row (x1 + 2 x2 = 1.25), so G = [5];
    the other rows have |row| sums 1 and column sums (1, 1), so beta =
    (1 + 1e-12) * 1 * 1; the metric norm of b is 1.25^2 / 5 plus the other
    rows' squared bounds (1 + 1) over beta; sigma = that norm / ||c||_2 and
    lambda_A = 2 beta."""
    qp, _ = Q.gram_fixture_qp("mixed")
    o = Q.OracleSession(qp, gram_mode=2, use_scaling=0, print_frequency=0)
    beta = (1 + 1e-12) * 1.0 * 1.0
    assert o.scalar("gram_active") == 1 and o.scalar("gram_rows") == 1
    assert o.scalar("gram_beta") == beta
    assert np.array_equal(o.read("gram_G"), [5.0])
    normb = math.sqrt(1.25 ** 2 / 5.0 + 2.0 / beta)
    c = -np.array([[2.0, 0.3], [0.3, 1.0]]) @ [0.25, 0.5]
    assert abs(o.scalar("sigma") - normb / np.linalg.norm(c)) <= 1e-15 * normb / np.linalg.norm(c)
    assert o.scalar("lambda_max_A") == 2 * beta
    assert o.scalar("noC") == 0


def _dense(ptr, ind, val, nrow, ncol):
    M = np.zeros((nrow, ncol))
    for r in range(nrow):
        for k in range(int(ptr[r]), int(ptr[r + 1])):
            M[r, int(ind[k])] += val[k]
    return M


def _gram_map(o):
    L = Q.oracle_lib()
    m = o.m
    out = (Q.C.c_long * max(m, 1))()
    assert L.or_session_gram_map(o.h, out, m) == m
    return np.array(out[:m])


def _audit(o, bodies):
    """Advance `bodies` loop bodies, take one more iteration step by hand and
    check it against dense linear algebra (tests/polishing/test_gpu.py:20-36)."""
    n, m = o.n, o.m
    A = _dense(*o.matrix("A"), m, n)
    gmap = _gram_map(o)
    eq = np.flatnonzero(gmap >= 0)
    rest = np.flatnonzero(gmap < 0)
    assert np.array_equal(gmap[eq], np.arange(len(eq)))
    e = len(eq)
    G = o.read("gram_G").reshape(e, e)
    EEt = A[eq] @ A[eq].T
    assert np.max(np.abs(G - EEt)) <= 1e-13 * (1 + np.max(np.abs(EEt)))
    beta = o.scalar("gram_beta")
    defects = []
    for k in bodies:
        st, fin = o.advance(k)
        assert st == 0 and not fin
        y_old = o.read("y")
        inner = int(o.scalar("inner"))
        assert o.step(inner, 1) == 0
        sigma = o.scalar("sigma")
        r, step = o.read("gram_r"), o.read("gram_step")
        assert np.array_equal(r, o.read("Ax")[eq] - o.read("gram_b"))
        want = np.linalg.solve(G, r)
        defects.append(np.max(np.abs(step - want)) / (1 + np.max(np.abs(want))))
        yb = o.read("y_bar")
        ye = y_old.copy()
        ye[eq] -= step / (2 * sigma)
        ye[rest] = (o.read("s")[rest] - o.read("Rp")[rest]) / (2 * sigma * beta)
        defects.append(np.max(np.abs(yb - ye)) / (1 + np.max(np.abs(ye))))
        dy = o.read("dy")
        qv = 2 * dy[eq] @ G @ dy[eq] + 2 * beta * (dy[rest] @ dy[rest])
        got = Q.oracle_lib().or_session_gram_quadratic(o.h)
        defects.append(abs(got - qv) / (1 + abs(qv)))
    return max(defects)


@pytest.mark.parametrize("kind", ["affine", "mixed"])
def test_gram_trace_audit_reference_fixtures(kind):
    qp, _ = Q.gram_fixture_qp(kind)
    o = Q.OracleSession(qp, gram_mode=2, print_frequency=0)
    assert _audit(o, [0, 1, 5, 40]) < 1e-10


def test_gram_trace_audit_random_model():
    qp = Q.random_qp(seed=11, n=120, m_eq=40, m_in=30, a_row_nnz=6, b_rows=30, b_row_nnz=4)
    o = Q.OracleSession(qp, gram_mode=2, print_frequency=0)
    assert o.scalar("gram_active") == 1 and o.scalar("gram_rows") == 40
    assert _audit(o, [0, 3, 30, 100]) < 1e-10


@pytest.mark.parametrize("g", [3, 6])
def test_gram_trace_audit_grid_model(g):
    qp = Q.gram_fixture_spatial(g)
    o = Q.OracleSession(qp, print_frequency=0)  # AUTO
    assert o.scalar("gram_active") == 2 and o.scalar("gram_rows") == g * g
    assert _audit(o, [0, 2, 50]) < 1e-10


def test_gram_audit_catches_planted_error():
    """A metric step computed with G off by one percent in one entry fails
    the audit's tolerance by orders of magnitude."""
    qp = Q.random_qp(seed=11, n=120, m_eq=40, m_in=30, a_row_nnz=6, b_rows=30, b_row_nnz=4)
    o = Q.OracleSession(qp, gram_mode=2, print_frequency=0)
    o.advance(5)
    o.step(int(o.scalar("inner")), 1)
    e = int(o.scalar("gram_rows"))
    G = o.read("gram_G").reshape(e, e).copy()
    G[3, 3] *= 1.01
    want = np.linalg.solve(G, o.read("gram_r"))
    step = o.read("gram_step")
    assert np.max(np.abs(step - want)) / (1 + np.max(np.abs(want))) > 1e-6


@pytest.mark.parametrize("g", [3, 8, 15])
def test_gram_auto_grid_model_reaches_the_unmetric_optimum(g):
    qp = Q.gram_fixture_spatial(g)
    o_off, r_off, _ = _solve(qp, gram_mode=0)
    o_on, r_on, _ = _solve(qp)
    assert o_on.scalar("gram_active") == 2 and o_off.scalar("gram_active") == 0
    for r in (r_off, r_on):
        assert Q.STATUS[r.status] == "OPTIMAL" and r.residuals <= 1e-8
    assert abs(r_on.primal_obj - r_off.primal_obj) <= 1e-7 * (1 + abs(r_off.primal_obj))


def test_gram_auto_rejects_perturbed_grid():
    """The AUTO rule wants the exact stencil and a homogeneous prefix: a
    state coefficient of 3.5 instead of 4, a neighbour coupling of -0.5
    instead of -1, or a nonzero bound on a stencil row each turn it off
    (gram_policy.cpp:11-47)."""
    base = Q.gram_fixture_spatial(4)
    o = Q.OracleSession(base, print_frequency=0)
    assert o.scalar("gram_mode") == 2
    bad = Q.gram_fixture_spatial(4)
    k = int(np.flatnonzero(bad.a_val == 4.0)[0])
    bad.a_val = bad.a_val.copy()
    bad.a_val[k] = 3.5
    assert Q.OracleSession(bad, print_frequency=0).scalar("gram_mode") == 0
    bad = Q.gram_fixture_spatial(4)
    kc = 4 * 4  # state columns start after the 4g controls
    cols = bad.a_ind.astype(np.int64)
    k = int(np.flatnonzero((bad.a_val == -1.0) & (cols >= kc) & (cols < kc + 16)
                           & (np.arange(len(cols)) < int(bad.a_ptr[16])))[0])
    bad.a_val = bad.a_val.copy()
    bad.a_val[k] = -0.5
    assert Q.OracleSession(bad, print_frequency=0).scalar("gram_mode") == 0
    bad = Q.gram_fixture_spatial(4)
    bad.AL, bad.AU = bad.AL.copy(), bad.AU.copy()
    bad.AL[0] = bad.AU[0] = 0.5
    o = Q.OracleSession(bad, print_frequency=0)
    assert o.status == 0 and o.scalar("gram_mode") == 0


def test_gram_random_models_reach_the_unmetric_optimum():
    for seed in range(4):
        qp = Q.random_qp(seed=seed, n=200, m_eq=60, m_in=40, a_row_nnz=8, b_rows=50, b_row_nnz=4)
        o_off, r_off, _ = _solve(qp, gram_mode=0)
        o_on, r_on, _ = _solve(qp, gram_mode=2)
        assert o_on.scalar("gram_active") == 1
        assert Q.STATUS[r_on.status] == "OPTIMAL" and Q.STATUS[r_off.status] == "OPTIMAL"
        assert abs(r_on.primal_obj - r_off.primal_obj) <= 1e-7 * (1 + abs(r_off.primal_obj))
