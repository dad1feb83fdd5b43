"""The oracle (oracle/hprqp_oracle.c) pinned before it is trusted:

* against the reference's own outputs -- golden traces written by
  tests/golden/make_golden.py from the reference built from source -- bitwise;
* against the live reference build when oracle/_ref is present -- bitwise;
* against the closed forms the reference's unit tests assert
  (tests/cpu/test_workspace.c, test_eigs.c, test_iteration.c,
  test_residuals.c, test_sparse_helpers.c, test_builder_preprocess.c);
* against mathematics: a brute-force active-set KKT solve on small QPs, the
  largest eigenvalue from numpy, and the Moreau decomposition for the cone
  projection.
"""
import itertools
import math
from pathlib import Path

import numpy as np
import pytest

import qp_libs as Q

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "reference_traces.npz")
KS = [1, 2, 3, 5, 12, 50, 99, 100, 101, 150, 300]
RANDOM = {"rand60": dict(seed=1, n=60, m_eq=20, m_in=10, a_row_nnz=5, b_rows=15, b_row_nnz=3),
          "rand200": dict(seed=2, n=200, m_eq=60, m_in=40, a_row_nnz=8, heavy_rows=2, heavy_nnz=100,
                          b_rows=50, b_row_nnz=4)}


def all_cases():
    out = dict(Q.reference_fixtures())
    for k, kw in RANDOM.items():
        out[k] = kw
    return out


def oracle_run(qp, k):
    o = Q.OracleSession(qp, **({} if k is None else {"max_iter": k}))
    assert o.status == 0
    return o.solve()


@pytest.mark.parametrize("name", list(all_cases().keys()))
def test_oracle_equals_reference_golden_traces(name):
    qp = all_cases()[name]
    if isinstance(qp, dict):
        qp = Q.random_qp(**qp)
    for k in KS + [None]:
        tag = f"{name}/{'full' if k is None else k}"
        if tag + "/scalars" not in GOLD.files:
            continue
        st, it, obj, kkt, gap = GOLD[tag + "/scalars"]
        r, x, y, z, w = oracle_run(qp, k)
        assert (r.status, r.iter) == (int(st), int(it)), tag
        assert (r.primal_obj, r.residuals, r.gap) == (obj, kkt, gap), tag
        for v, arr in zip("xyzw", (x, y, z, w)):
            assert np.array_equal(arr, GOLD[tag + "/" + v]), (tag, v)


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("kw", [dict(seed=0), dict(seed=5, n=500, m_eq=150, m_in=100, a_row_nnz=6,
                                                   heavy_rows=3, heavy_nnz=300, b_rows=120, b_row_nnz=5)])
def test_oracle_equals_live_reference(kw):
    qp = Q.random_qp(**kw)
    for k in (7, 100, 333, None):
        r = Q.binding_solve(Q.ref_lib(), qp, use_gpu=0, **({} if k is None else {"max_iter": k}))
        o, x, y, z, w = oracle_run(qp, k)
        assert Q.STATUS[o.status] == r["status"] and o.iter == r["iter"]
        assert o.primal_obj == r["objective"] and o.residuals == r["residual"]
        for v, arr in zip("xyzw", (x, y, z, w)):
            assert np.array_equal(arr, r[v]), v


def test_workspace_closed_forms():
    # tests/cpu/test_workspace.c:20-37: sigma0 = 2/sqrt(5), lambda_A = 2.1, lambda_Q = 4.0
    o = Q.OracleSession(Q.reference_fixtures()["STANDARD"], use_scaling=0)
    s = 2.0 / math.sqrt(5.0)
    assert abs(o.scalar("sigma") - s) <= 1e-12
    assert abs(o.scalar("lambda_max_A") - 2.1) <= 1e-12
    assert abs(o.scalar("lambda_max_Q") - 4.0) <= 1e-12
    assert o.scalar("Q_is_diag") == 1.0
    assert np.allclose(o.read("fact"), [s / (1 + 2 * s), s / (1 + 4 * s)], atol=1e-12, rtol=0)
    assert np.allclose(o.read("fact_M"), [s * s / (1 + 2 * s), s * s / (1 + 4 * s)], atol=1e-12, rtol=0)


def test_eigs_closed_forms():
    # tests/cpu/test_eigs.c:19-20: lambda_A = 5.25 +- 1e-8, lambda_Q = 3.15 +- 1e-3
    o = Q.OracleSession(Q.reference_fixtures()["DENSE_Q"], use_scaling=0)
    assert abs(o.scalar("lambda_max_A") - 5.25) <= 1e-8
    assert abs(o.scalar("lambda_max_Q") - 3.15) <= 1e-3


def test_power_iteration_matches_numpy_eigenvalue():
    qp = Q.random_qp(seed=4, n=80, m_eq=30, m_in=10, a_row_nnz=6, b_rows=20, b_row_nnz=4)
    o = Q.OracleSession(qp)
    A = np.zeros((o.m, o.n))
    rp, ci, v = o.matrix("A")
    for r in range(o.m):
        A[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    lam = np.linalg.eigvalsh(A @ A.T).max()
    # power iteration stops at relative residual 1e-4; the estimate is scaled by 1.05
    assert abs(o.scalar("lambda_max_A") / 1.05 - lam) <= 2e-3 * lam


def test_iteration_step_closed_form():
    # tests/cpu/test_iteration.c:53-103: one HPR step on STANDARD without scaling
    o = Q.OracleSession(Q.reference_fixtures()["STANDARD"], use_scaling=0)
    assert o.step(0, 1) == 0
    s = 2.0 / math.sqrt(5.0)
    yb = 2.0 / (2.1 * s)
    w0, w1 = s / (1 + 2 * s) * yb, s / (1 + 4 * s) * yb
    assert np.allclose(o.read("x_bar"), [0, 0], atol=1e-12)
    assert np.allclose(o.read("z_bar"), [2, 1], atol=1e-12)
    assert np.allclose(o.read("y_bar"), [yb], atol=1e-12) and np.allclose(o.read("y"), [yb], atol=1e-12)
    assert np.allclose(o.read("s"), [2.0], atol=1e-12)
    assert np.allclose(o.read("w_bar"), [w0, w1], atol=1e-12) and np.allclose(o.read("w"), [w0, w1], atol=1e-12)
    assert np.allclose(o.read("ATy"), [yb, yb], atol=1e-12) and np.allclose(o.read("ATdy"), [yb, yb], atol=1e-12)
    # to_check = 0 leaves the diagnostic vectors alone (test_iteration.c:105-145)
    o2 = Q.OracleSession(Q.reference_fixtures()["STANDARD"], use_scaling=0)
    for name in ("dx", "dw", "ATdy"):
        o2.write(name, [-7.0, -7.0])
    o2.write("dy", [-7.0])
    o2.step(0, 0)
    for name in ("dx", "dw", "ATdy", "dy"):
        assert np.all(o2.read(name) == -7.0), name


def test_residuals_closed_form():
    # tests/cpu/test_residuals.c:24-53 (CPU dual-objective convention)
    o = Q.OracleSession(Q.reference_fixtures()["RESIDUALS"], use_scaling=0)
    o.write("x_bar", [1.0, 1.0])
    o.write("y_bar", [0.5])
    o.write("z_bar", [0.2, 0.3])
    o.write("s", [0.25])
    r = o.residuals(1)
    assert abs(r["pobj"] - 3.0) <= 1e-12 and abs(r["dobj"] + 2.375) <= 1e-12
    assert abs(r["err_rd"] - 0.46) <= 1e-12 and abs(r["err_rp"]) <= 1e-12
    assert abs(r["gap"] - 1.34375) <= 1e-12 and abs(r["kkt"] - 1.34375) <= 1e-12
    o.write("x_bar", [11.0, 1.0])
    o.write("y_bar", [0.0])
    o.write("z_bar", [0.0, 0.0])
    o.write("s", [0.0])
    assert abs(o.residuals(0)["err_rp"] - 1.0) <= 1e-12


def test_sparse_known_answers():
    # tests/test_sparse_helpers.c:60,73: A x = [13, 19]; transpose gives A^T y = [13, 38, 147]
    L = Q.oracle_lib()
    rp = np.array([0, 2, 4], np.uint64)
    ci = np.array([0, 2, 1, 2], np.uint64)
    v = np.array([1.0, 4.0, 2.0, 5.0])
    x = np.array([1.0, 2.0, 3.0])
    y = np.zeros(2)
    L.or_csr_matvec(2, rp.ctypes.data_as(Q.c_size_p), ci.ctypes.data_as(Q.c_size_p), v.ctypes.data_as(Q.c_double_p),
                    x.ctypes.data_as(Q.c_double_p), y.ctypes.data_as(Q.c_double_p))
    assert list(y) == [13.0, 19.0]
    trp, tci, tv = np.zeros(4, np.uint64), np.zeros(4, np.uint64), np.zeros(4)
    L.or_csr_transpose(2, 3, rp.ctypes.data_as(Q.c_size_p), ci.ctypes.data_as(Q.c_size_p),
                       v.ctypes.data_as(Q.c_double_p), trp.ctypes.data_as(Q.c_size_p),
                       tci.ctypes.data_as(Q.c_size_p), tv.ctypes.data_as(Q.c_double_p))
    yt = np.zeros(3)
    L.or_csr_matvec(3, trp.ctypes.data_as(Q.c_size_p), tci.ctypes.data_as(Q.c_size_p),
                    tv.ctypes.data_as(Q.c_double_p), y.ctypes.data_as(Q.c_double_p), yt.ctypes.data_as(Q.c_double_p))
    assert list(yt) == [13.0, 38.0, 147.0]


def test_transpose_is_stable_sort_by_column():
    qp = Q.random_qp(seed=9, n=40, m_eq=30, m_in=30, a_row_nnz=7, b_rows=5, b_row_nnz=3)
    o = Q.OracleSession(qp, use_scaling=0)
    rp, ci, v = o.matrix("A")
    trp, tci, tv = o.matrix("AT")
    rows = np.repeat(np.arange(o.m), np.diff(rp.astype(np.int64)))
    order = np.lexsort((rows, ci))            # by column, then row
    assert np.array_equal(tci, rows[order]) and np.array_equal(tv, v[order])


def test_preprocess_known_answers():
    # tests/test_builder_preprocess.c:22-84: Q = [2 6; 0 4] -> [2 3; 3 4]; empty and free rows deleted
    inf = math.inf
    qp = Q.QP(2, 3, np.array([0, 2, 3], np.uint64), np.array([0, 1, 1], np.uint64), np.array([2.0, 6.0, 4.0]),
              np.array([0, 2, 2, 2], np.uint64), np.array([0, 1], np.uint64), np.array([1.0, -1.0]),
              np.zeros(2), np.array([0.0, -inf, 1.0]), np.array([3.0, inf, 2.0]), np.array([-2.0, -2.0]),
              np.array([2.0, 2.0]))
    o = Q.OracleSession(qp, use_scaling=0)
    assert o.m == 1
    rp, ci, v = o.matrix("Q")
    assert list(rp) == [0, 2, 4] and list(ci) == [0, 1, 0, 1] and list(v) == [2.0, 3.0, 3.0, 4.0]
    assert list(o.read("AL")) == [0.0] and list(o.read("AU")) == [3.0]


def brute_force_qp(Qd, c, A, AL, AU, l, u):
    """Minimise 1/2 x'Qx + c'x over box and rows by enumerating active sets."""
    n, m = len(c), A.shape[0]
    rows = np.vstack([A, np.eye(n)])
    lo = np.concatenate([AL, l])
    hi = np.concatenate([AU, u])
    best, best_x = math.inf, None
    for choice in itertools.product((0, 1, 2), repeat=m + n):  # 0 free, 1 at lower, 2 at upper
        act = [(i, lo[i] if ch == 1 else hi[i]) for i, ch in enumerate(choice) if ch]
        if any(not math.isfinite(b) for _, b in act):
            continue
        k = len(act)
        G = np.zeros((n + k, n + k))
        G[:n, :n] = Qd
        rhs = np.zeros(n + k)
        rhs[:n] = -c
        for j, (i, b) in enumerate(act):
            G[:n, n + j] = rows[i]
            G[n + j, :n] = rows[i]
            rhs[n + j] = b
        try:
            sol = np.linalg.solve(G, rhs)
        except np.linalg.LinAlgError:
            continue
        x = sol[:n]
        if np.all(rows @ x >= lo - 1e-9) and np.all(rows @ x <= hi + 1e-9):
            f = 0.5 * x @ Qd @ x + c @ x
            if f < best - 1e-12:
                best, best_x = f, x
    return best, best_x


@pytest.mark.parametrize("name", ["STANDARD", "SCALE", "RESIDUALS", "CONTROL_NONDIAG", "GPU_ITERATION_STANDARD"])
def test_oracle_solution_matches_brute_force_kkt(name):
    qp = Q.reference_fixtures()[name]
    r, x, y, z, w = oracle_run(qp, None)
    assert Q.STATUS[r.status] == "OPTIMAL"
    f, xs = brute_force_qp(qp.dense_Q(), qp.c, qp.dense_A(), qp.AL, qp.AU, qp.l, qp.u)
    assert abs(r.primal_obj - f) <= 1e-6 * max(1.0, abs(f))
    assert np.max(np.abs(x - xs)) <= 1e-5


def test_cone_projection_against_closed_form_and_moreau():
    L = Q.oracle_lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        d = int(rng.integers(2, 12))
        v = rng.normal(size=d) * 10.0 ** rng.integers(-3, 4)
        p = np.zeros(d)
        L.or_soc_project(p.ctypes.data_as(Q.c_double_p), v.ctypes.data_as(Q.c_double_p), d)
        t, xv = v[0], v[1:]
        nv = np.linalg.norm(xv)
        if nv <= t:
            want = v
        elif nv <= -t:
            want = np.zeros(d)
        else:
            want = np.concatenate([[(nv + t) / 2], (nv + t) / (2 * nv) * xv])
        assert np.allclose(p, want, rtol=1e-13, atol=1e-13 * np.abs(v).max())
        # Moreau: v = P_K(v) - P_K(-v) for the self-dual Lorentz cone
        q = np.zeros(d)
        mv = -v
        L.or_soc_project(q.ctypes.data_as(Q.c_double_p), mv.ctypes.data_as(Q.c_double_p), d)
        assert np.allclose(p - q, v, rtol=1e-12, atol=1e-12 * np.abs(v).max())
        assert abs(p @ q) <= 1e-9 * max(1.0, np.abs(v).max() ** 2)


def test_sigma_search_sensitivity():
    """Why iterates after the first restarts are compared loosely: with lambda_A
    changed by one ulp the oracle's own trajectory agrees to ~1e-15 through
    iteration 150 but sigma differs by ~1e-8 relative after the second
    restart, because the golden-section search (control.c:45-77) resolves the
    minimiser only to about sqrt(eps)."""
    qp = Q.random_qp(seed=1, n=60, m_eq=20, m_in=10, a_row_nnz=5, b_rows=15, b_row_nnz=3)
    base = Q.OracleSession(qp)
    la, lq = base.scalar("lambda_max_A"), base.scalar("lambda_max_Q")

    def run(k, lam):
        o = Q.OracleSession(qp, max_iter=k, lambda_A_given=lam, lambda_Q_given=lq)
        r = o.solve()
        return r[1], o.scalar("sigma")
    x1, _ = run(150, la)
    x2, _ = run(150, la * (1 + 2.0 ** -52))
    assert np.max(np.abs(x1 - x2)) <= 1e-13 * np.max(np.abs(x1))
    _, s1 = run(250, la)
    _, s2 = run(250, la * (1 + 2.0 ** -52))
    assert 1e-12 < abs(s1 - s2) / s1 < 1e-6
