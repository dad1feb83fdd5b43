"""Cone (SOC) part of the oracle. The reference has no CPU path for cones
(src/shared/solver.c:250-253), so the restatement of the GPU semantics
(src/gpu/soc_*.cuh) is pinned by the reference's own analytic SOC cases and
certificate (tests/gpu/test_gpu_soc.cu:152-273): Euclidean projection of
(0,3,4) onto the cone -> (2.5,1.5,2); min t with (t-2,3,4) in the cone -> 7;
the constant affine cone (feasible -> x=1, infeasible -> MAX_ITER with
residual >= 0.49); and the mixed 8x7 model whose optimum is constructed
(x = (0.25,-0.5,5,3,4,2,0,0)).
"""
import numpy as np
import pytest

import qp_libs as Q


@pytest.mark.parametrize("scaling", [0, 1])
@pytest.mark.parametrize("name", ["proj_cone", "affine_cone", "constant_cone", "constant_cone_infeasible"])
def test_soc_simple_models(name, scaling):
    qp, want = Q.soc_simple_models()[name]
    kw = dict(use_scaling=scaling, stoptol=1e-9 if name != "constant_cone" else 1e-8, print_frequency=0,
              max_iter=100000 if name != "constant_cone_infeasible" and name != "constant_cone" else 1000)
    o = Q.OracleSession(qp, **kw)
    assert o.status == 0
    r, x, y, z, w = o.solve()
    if want is None:
        assert Q.STATUS[r.status] == "MAX_ITER" and r.residuals >= 0.49
        return
    assert Q.STATUS[r.status] == "OPTIMAL", (name, r.status, r.iter, r.residuals)
    assert np.max(np.abs(x - want)) <= 2e-7 if name != "constant_cone" else 1e-7
    cert = Q.soc_certificate(qp, x, y, z)
    assert max(cert.values()) < 2e-6, cert


@pytest.mark.parametrize("q_kind", [0, 1, 2])
@pytest.mark.parametrize("scaling", [0, 1])
def test_soc_mixed_model(q_kind, scaling):
    qp = Q.soc_mixed_model(q_kind)
    o = Q.OracleSession(qp, use_scaling=scaling, stoptol=1e-9, print_frequency=0, max_iter=100000, time_limit=30.0)
    r, x, y, z, w = o.solve()
    assert Q.STATUS[r.status] == "OPTIMAL", (r.status, r.iter, r.residuals)
    cert = Q.soc_certificate(qp, x, y, z)
    assert max(cert.values()) < 2e-6, cert
    if q_kind:
        assert np.max(np.abs(x - [0.25, -0.5, 5, 3, 4, 2, 0, 0])) <= 2e-5


def _ld_product(ptr, ind, val, v):
    """Row sums in extended precision, as the reference's host restatement
    forms them (tests/gpu/test_gpu_soc.cu:28-38)."""
    out = np.zeros(len(ptr) - 1)
    v = np.asarray(v, np.longdouble)
    for i in range(len(ptr) - 1):
        a, b = int(ptr[i]), int(ptr[i + 1])
        out[i] = float(np.sum(np.asarray(val[a:b], np.longdouble) * v[ind[a:b].astype(np.int64)],
                              dtype=np.longdouble))
    return out


def _ld_projection(x, first, end):
    """Lorentz projection of x[first:end] with an extended-precision norm
    built by repeated hypot, as in tests/gpu/test_gpu_soc.cu:40-53."""
    x = np.array(x, float)
    t = np.longdouble(x[first])
    norm = np.longdouble(0)
    for j in range(first + 1, end):
        norm = np.hypot(norm, np.longdouble(x[j]))
    if norm <= t:
        return x
    if norm <= -t:
        x[first:end] = 0.0
        return x
    x[first] = float((norm + t) / 2)
    for j in range(first + 1, end):
        x[j] = float((1 + t / norm) * np.longdouble(x[j]) / 2)
    return x


@pytest.mark.parametrize("to_check", [0, 1])
@pytest.mark.parametrize("q_kind", [0, 2])
def test_soc_single_step_matches_host_restatement(q_kind, to_check):
    """One cone step of the oracle against the reference's own host
    restatement of it (tests/gpu/test_gpu_soc.cu:275-328): the mixed model
    unscaled, state x_j = 0.1(j+1), w_j = -0.03(j+2), y_i = 0.04(i+1), inner
    iteration 4 (Halpern weight 1/6), every updated vector within 3e-13 of
    1 + its infinity norm."""
    qp = Q.soc_mixed_model(q_kind)
    o = Q.OracleSession(qp, use_scaling=0, print_frequency=0)
    assert o.status == 0
    assert (q_kind == 0 or o.scalar("Q_is_diag") == 0) and o.scalar("noC") == 0
    n, m = o.n, o.m
    Qp, Qi, Qv = o.matrix("Q")
    Ap, Ai, Av = o.matrix("A")
    Tp, Ti, Tv = o.matrix("AT")
    c, l, u, AL, AU = (o.read(k) for k in ("c", "l", "u", "AL", "AU"))
    rhs = qp.soc_rhs
    x = 0.1 * np.arange(1, n + 1)
    oldw = -0.03 * np.arange(2, n + 2)
    y = 0.04 * np.arange(1, m + 1)
    aty = _ld_product(Tp, Ti, Tv, y)
    qw = _ld_product(Qp, Qi, Qv, oldw)
    sigma, lam_q, lam_a = o.scalar("sigma"), o.scalar("lambda_max_Q"), o.scalar("lambda_max_A")
    f2 = 1 / (1 + sigma * lam_q)
    tau = sigma * lam_a
    a = 1.0 / 6
    b = 1 - a
    raw = x + sigma * (-qw + aty - c)
    xb = raw.copy()
    xb[:2] = np.minimum(np.maximum(raw[:2], l[:2]), u[:2])
    xb = _ld_projection(xb, 2, 5)
    xb = _ld_projection(xb, 5, 8)
    xh = 2 * xb - x
    wb1 = (1 - f2) * oldw + f2 * xh
    ez = (xb - raw) / sigma
    temp = xh + sigma * (qw - _ld_product(Qp, Qi, Qv, wb1))
    ax = _ld_product(Ap, Ai, Av, temp)
    s = ax - rhs - tau * y
    rowraw = s.copy()
    s[:2] = np.minimum(np.maximum(s[:2], AL[:2]), AU[:2])
    s = _ld_projection(s, 2, 4)
    s = _ld_projection(s, 4, 7)
    yb = (s - rowraw) / tau
    yn = a * y + b * (2 * yb - y)
    wb = wb1 + sigma * f2 * (_ld_product(Tp, Ti, Tv, yb) - aty)
    wn = a * oldw + b * (2 * wb - oldw)
    xn = a * x + b * xh

    for name, v in (("x", x), ("last_x", x), ("w", oldw), ("last_w", oldw), ("y", y), ("last_y", y),
                    ("ATy", aty), ("last_ATy", aty)):
        assert o.write(name, v) == 0
    assert o.step(4, to_check) == 0
    expect = {"x_bar": xb, "z_bar": ez, "w_bar": wb, "y_bar": yb, "s": s, "x": xn, "y": yn, "w": wn}
    for name, e in expect.items():
        got = o.read(name)
        err = np.max(np.abs(got - e)) / (1 + np.max(np.abs(e)))
        assert err <= 3e-13, (name, err)


def test_soc_single_step_detects_planted_sign_error():
    """The restatement above is sharp enough to catch a sign slip: flipping
    the sign of the A'y term in the raw x point moves x_bar far beyond 3e-13."""
    qp = Q.soc_mixed_model(2)
    o = Q.OracleSession(qp, use_scaling=0, print_frequency=0)
    n, m = o.n, o.m
    Tp, Ti, Tv = o.matrix("AT")
    y = 0.04 * np.arange(1, m + 1)
    aty = _ld_product(Tp, Ti, Tv, y)
    x = 0.1 * np.arange(1, n + 1)
    for name, v in (("x", x), ("last_x", x), ("y", y), ("last_y", y), ("ATy", -aty), ("last_ATy", -aty)):
        assert o.write(name, v) == 0
    assert o.step(4, 1) == 0
    bad = o.read("x_bar")
    o2 = Q.OracleSession(qp, use_scaling=0, print_frequency=0)
    for name, v in (("x", x), ("last_x", x), ("y", y), ("last_y", y), ("ATy", aty), ("last_ATy", aty)):
        assert o2.write(name, v) == 0
    assert o2.step(4, 1) == 0
    assert np.max(np.abs(bad - o2.read("x_bar"))) > 1e-3
