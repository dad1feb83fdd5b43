"""Batched solver: every instance of a batch matches a plain solve of the
same instance (status, objective, iterate) and the oracle."""
import math

import numpy as np
import pytest

import qp_libs as Q
from test_gpu_parity import build

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    return hprqp_amd


def test_batch_matches_plain_and_oracle(H):
    qps = Q.batched_instances(count=24, seed=11, n_lo=20, n_hi=400)
    models = [build(H, qp) for qp in qps]
    print("batch of", len(models), flush=True)
    res, secs = H.solve_batch(models, H.Parameters(verbose=False, stoptol=1e-6), bodies_per_launch=50)
    for k, (qp, model, r) in enumerate(zip(qps, models, res)):
        plain = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-6))
        o = Q.OracleSession(qp, stoptol=1e-6)
        orr, x, y, z, w = o.solve()
        assert r.status == plain.status == "OPTIMAL" == Q.STATUS[orr.status], (k, r.status, plain.status)
        assert abs(r.iter - orr.iter) <= 10, (k, r.iter, orr.iter)
        assert abs(r.objective - orr.primal_obj) <= 1e-6 * max(1.0, abs(orr.primal_obj)), k
        assert np.max(np.abs(r.x - x)) <= 1e-5 * max(1.0, np.max(np.abs(x))), k


def test_batch_mixed_dense_and_csr_instances(H, monkeypatch):
    """One batch holding dense-form instances (full Q and A: dense storage is
    no larger than CSR) next to CSR instances, so that every chunk launches
    both kernels, with the CSR ones in the tail kernel (HPRQP_BATCH_LONE=1)
    and in the shared kernel (=0): each instance as its plain solve."""
    Qd, c, A, AL, AU, l, u = Q.layer_instances(3, n=40, m_eq=10, m_in=20, seed=4)
    models = [H.build_from_qabc(Qd[k], c[k], A[k], AL[k], AU[k], l[k], u[k]) for k in range(3)]
    qps = Q.batched_instances(count=5, seed=17, n_lo=30, n_hi=500)
    models += [build(H, qp) for qp in qps]
    params = H.Parameters(verbose=False, stoptol=1e-8)
    plain = [H.optimize(mo, params) for mo in models]
    for lone in ("1", "0"):
        monkeypatch.setenv("HPRQP_BATCH_LONE", lone)
        print("mixed batch, lone", lone, flush=True)
        res, _ = H.solve_batch(models, params, bodies_per_launch=30)
        for k, (r, p_) in enumerate(zip(res, plain)):
            assert r.status == p_.status == "OPTIMAL", (lone, k, r.status, p_.status)
            assert abs(r.objective - p_.objective) <= 1e-7 * max(1.0, abs(p_.objective)), (lone, k)


def test_batch_ragged_edge_instances(H):
    """One batch holding a one-variable QP, an LP (Q = 0), an unconstrained
    QP (m = 0) and a generated QP: each instance as the plain solve."""
    inf = math.inf
    qp = Q.random_qp(seed=5, n=50, m_eq=10, m_in=8, a_row_nnz=4, b_rows=12, b_row_nnz=3)
    models = [
        H.build_from_qabc([[2.0]], [1.0], [[1.0]], [0.0], [1.0], [-10.0], [10.0]),
        H.build_from_qabc([[0.0, 0.0], [0.0, 0.0]], [1.0, 2.0], [[1.0, 1.0]], [1.0], [inf], [0.0, 0.0],
                          [10.0, 10.0]),
        H.build_from_qabc([[2.0, 0.0], [0.0, 2.0]], [-4.0, -2.0], [], [], [], [-inf, -inf], [inf, inf]),
        build(H, qp),
    ]
    params = H.Parameters(verbose=False, stoptol=1e-8)
    res, _ = H.solve_batch(models, params, bodies_per_launch=25)
    expect = [0.0, 1.0, -5.0, None]
    for k, (model, r) in enumerate(zip(models, res)):
        plain = H.optimize(model, params)
        assert r.status == plain.status == "OPTIMAL", (k, r.status, plain.status)
        assert abs(r.objective - plain.objective) <= 1e-7 * max(1.0, abs(plain.objective)), k
        assert np.max(np.abs(r.x - plain.x)) <= 1e-5 * max(1.0, np.max(np.abs(plain.x))), k
        assert len(r.y) == len(plain.y), k
        if expect[k] is not None:
            assert r.objective == pytest.approx(expect[k], abs=1e-6), (k, r.objective)


def test_batch_fixture_traces(H):
    """Reference fixtures (tests/test_support.c) as one batch at max_iter
    100: each instance's iterate equals the reference's golden trace."""
    gold = np.load(Q.ROOT / "tests" / "golden" / "reference_traces.npz")
    names = ["STANDARD", "SCALE", "RESIDUALS", "DENSE_Q", "CONTROL_NONDIAG", "TALL_DIAG", "GPU_ITERATION_STANDARD",
             "EMPTY_A"]
    names += ["STANDARD_NOC"]
    F = Q.reference_fixtures()
    models = [build(H, F[n]) for n in names]
    res, _ = H.solve_batch(models, H.Parameters(verbose=False, max_iter=100), bodies_per_launch=7)
    for n, r in zip(names, res):
        want = gold[f"{n}/100/scalars"]
        assert r.iter == int(want[1]), n
        assert Q.STATUS_CODE[r.status] == int(want[0]), (n, r.status)
        scale = max(1.0, float(np.max(np.abs(gold[f"{n}/100/x"]))))
        assert np.max(np.abs(r.x - gold[f"{n}/100/x"])) <= 1e-10 * scale, n


def _layer_models(H, count, n, m_eq, m_in, seed, boxed_every=2):
    Qd, c, A, AL, AU, l, u = Q.layer_instances(count, n=n, m_eq=m_eq, m_in=m_in, seed=seed)
    models = []
    for k in range(count):
        lk, uk = l[k], u[k]
        if k % boxed_every == 1:  # finite bounds: the general (not noC) iteration
            lk, uk = np.full(n, -2.0), np.full(n, 2.0)
        models.append(H.build_from_qabc(Qd[k], c[k], A[k], AL[k], AU[k], lk, uk))
    return models


def _batch_x(H, models, monkeypatch, dense, **params):
    monkeypatch.setenv("HPRQP_BATCH_DENSE", "1" if dense else "0")
    res, _ = H.solve_batch(models, H.Parameters(verbose=False, **params), bodies_per_launch=40)
    return res


@pytest.mark.parametrize("n,m_eq,m_in", [(48, 12, 40), (256, 64, 256)])
def test_batch_dense_mode_matches_sparse_mode(H, monkeypatch, n, m_eq, m_in):
    """Dense instances (column-form products over row-major Q, A and A^T)
    against the CSR engine on the same models. At n = 48 every product row
    is summed in the same sequential order in both modes; the modes differ
    in which thread owns which row, hence in the order of the workgroup
    reductions (dot products, norms), so iterates agree to rounding. At
    n = 256 the CSR rows (>= 64 entries) are also reduced by a wavefront
    (FMA and a tree)."""
    models = _layer_models(H, 6, n, m_eq, m_in, seed=3)
    print(f"dense vs sparse batch n={n}", flush=True)
    for max_iter in (1, 37, 100):
        rd = _batch_x(H, models, monkeypatch, True, max_iter=max_iter)
        rs = _batch_x(H, models, monkeypatch, False, max_iter=max_iter)
        for k, (a, b) in enumerate(zip(rd, rs)):
            assert a.iter == b.iter and a.status == b.status, (k, max_iter)
            for v in ("x", "y", "z"):
                va, vb = getattr(a, v), getattr(b, v)
                scale = max(1.0, float(np.max(np.abs(vb))))
                assert np.max(np.abs(va - vb)) <= 1e-10 * scale, (k, v, max_iter, float(np.max(np.abs(va - vb))))


def test_batch_dense_mixed_shapes_match_sparse_mode(H, monkeypatch):
    """Dense instances of different shapes in one batch (m below, at and
    above the 32-row tiles of the setup's G = A A^T, m < n and m > n) against
    the CSR engine on the same models, per iterate and to optimality."""
    shapes = [(10, 3, 4), (33, 8, 24), (70, 30, 35), (40, 16, 48), (20, 1, 2)]
    models = []
    for k, (n, meq, min_) in enumerate(shapes):
        models.append(_layer_models(H, 2, n, meq, min_, seed=20 + k)[k % 2])  # free and boxed variables in turn
    print("dense mixed shapes", flush=True)
    for max_iter in (1, 60):
        rd = _batch_x(H, models, monkeypatch, True, max_iter=max_iter)
        rs = _batch_x(H, models, monkeypatch, False, max_iter=max_iter)
        for k, (a, b) in enumerate(zip(rd, rs)):
            assert a.iter == b.iter and a.status == b.status, (k, max_iter)
            for v in ("x", "y", "z"):
                va, vb = getattr(a, v), getattr(b, v)
                scale = max(1.0, float(np.max(np.abs(vb))))
                assert np.max(np.abs(va - vb)) <= 1e-10 * scale, (k, v, max_iter)
    rd = _batch_x(H, models, monkeypatch, True, stoptol=1e-8)
    for k, (model, r) in enumerate(zip(models, rd)):
        plain = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
        assert r.status == plain.status == "OPTIMAL", (k, r.status, plain.status)
        assert abs(r.objective - plain.objective) <= 1e-7 * max(1.0, abs(plain.objective)), k


def test_batch_dense_mode_full_solves_match_plain(H, monkeypatch):
    models = _layer_models(H, 5, 96, 24, 96, seed=4)
    res = _batch_x(H, models, monkeypatch, True, stoptol=1e-8)
    for k, (model, r) in enumerate(zip(models, res)):
        plain = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
        assert r.status == plain.status == "OPTIMAL", (k, r.status, plain.status)
        assert abs(r.objective - plain.objective) <= 1e-7 * max(1.0, abs(plain.objective)), k
        assert r.residual <= 1e-8


def test_batch_from_dense_equals_model_batch(H, monkeypatch):
    """hprqp_hip_batch_create_dense (Q symmetrised on the device) against
    build_from_QAbc models batched in dense mode: identical iterates. Q is
    made unsymmetric, with entries at the 1e-15 drop tolerance, so the
    symmetrisation and its drop rule are exercised."""
    rng = np.random.default_rng(8)
    count, n, m_eq, m_in = 5, 40, 10, 30
    Qd, c, A, AL, AU, l, u = Q.layer_instances(count, n=n, m_eq=m_eq, m_in=m_in, seed=8)
    Qd = Qd + np.triu(rng.normal(scale=1e-3, size=(count, n, n)), 1)
    Qd[:, 0, 1] = 1.5e-15
    Qd[:, 1, 0] = 1e-16
    Qd[:, 2, 3] = 3e-15
    Qd[:, 3, 2] = -3e-15 + 1e-30
    l[1::2], u[1::2] = -2.0, 2.0
    monkeypatch.setenv("HPRQP_BATCH_DENSE", "1")
    models = [H.build_from_qabc(Qd[k], c[k], A[k], AL[k], AU[k], l[k], u[k]) for k in range(count)]
    for max_iter in (1, 60):
        params = H.Parameters(verbose=False, max_iter=max_iter)
        bm = H.Batch(models, params)
        bm.solve(25)
        bd = H.Batch.from_dense(Qd, c, A, AL, AU, l, u, params)
        bd.solve(25)
        sm, sd = bm.solution(), bd.solution()
        for key in ("x", "y", "z", "objective", "kkt"):
            np.testing.assert_array_equal(sd[key], sm[key], err_msg=f"{key} {max_iter}")
        assert sd["status"] == sm["status"] and (sd["iter"] == sm["iter"]).all()
        for k in range(count):  # bulk getter agrees with the per-instance one
            r = bm.result(k)
            np.testing.assert_array_equal(r.x, sm["x"][k * n:(k + 1) * n])


def test_batch_from_dense_rejects_rows_it_cannot_keep(H):
    Qd, c, A, AL, AU, l, u = Q.layer_instances(2, n=8, m_eq=2, m_in=3, seed=1)
    A[1, 2] = 0.0
    with pytest.raises(ValueError):
        H.Batch.from_dense(Qd, c, A, AL, AU, l, u)
    A, AL, AU = Q.layer_instances(2, n=8, m_eq=2, m_in=3, seed=1)[2:5]
    AL[0, 4], AU[0, 4] = -math.inf, math.inf
    with pytest.raises(ValueError):
        H.Batch.from_dense(Qd, c, A, AL, AU, l, u)


@pytest.mark.parametrize("lone", ["0", "1"])
def test_batch_graph_replays_equal_eager_launches(H, monkeypatch, lone):
    """The captured chunk graph, replayed many times (5 bodies per replay,
    host synchronisation and counter read-back between replays), gives the
    same bits as eager launches of the same kernels: the shared 256-thread
    kernel (HPRQP_BATCH_LONE=0) and the tail kernel (=1)."""
    qps = Q.batched_instances(count=12, seed=21, n_lo=30, n_hi=300)
    models = [build(H, qp) for qp in qps]
    monkeypatch.setenv("HPRQP_BATCH_WAVE_ROW", "32")
    monkeypatch.setenv("HPRQP_BATCH_LONE", lone)
    out = {}
    for graph in ("1", "0"):
        monkeypatch.setenv("HPRQP_BATCH_GRAPH", graph)
        print("graph", graph, flush=True)
        b = H.Batch(models, H.Parameters(verbose=False, stoptol=1e-6))
        assert b.wave_row == 32
        b.solve(5)
        out[graph] = b.solution()
    for key in ("x", "y", "z", "objective", "iter"):
        np.testing.assert_array_equal(out["1"][key], out["0"][key], err_msg=key)
    assert out["1"]["status"] == out["0"]["status"] == ["OPTIMAL"] * len(models)


def test_batch_lone_kernel_matches_shared_kernel(H, monkeypatch):
    """The tail kernel (one 1,024-thread workgroup per instance, four row
    blocks of a product in flight) against the shared 256-thread kernel on
    the same CSR instances: every row is summed alike, the workgroup
    reductions add their terms in another order, so iterates agree to
    rounding at fixed iteration counts, and both solve to the same optimum.
    HPRQP_BATCH_LONE=1 runs the tail kernel from the start, =0 never."""
    qps = Q.batched_instances(count=12, seed=31, n_lo=40, n_hi=2000)
    models = [build(H, qp) for qp in qps]
    out = {}
    for max_iter in (1, 37, 100, 100000):
        for lone in ("1", "0"):
            monkeypatch.setenv("HPRQP_BATCH_LONE", lone)
            print("lone", lone, "max_iter", max_iter, flush=True)
            b = H.Batch(models, H.Parameters(verbose=False, stoptol=1e-8, max_iter=max_iter))
            b.solve(50)
            out[lone] = b.solution()
        a, c = out["1"], out["0"]
        assert a["status"] == c["status"]
        if max_iter < 100000:
            assert (a["iter"] == c["iter"]).all()
            for v in ("x", "y", "z"):
                scale = max(1.0, float(np.max(np.abs(c[v]))))
                assert np.max(np.abs(a[v] - c[v])) <= 1e-10 * scale, (v, max_iter)
        else:
            assert c["status"] == ["OPTIMAL"] * len(models)
            rel = np.abs(a["objective"] - c["objective"]) / np.maximum(1.0, np.abs(c["objective"]))
            assert float(rel.max()) <= 1e-7, float(rel.max())


@pytest.mark.parametrize("wave_row", ["16", "1048576"])
def test_batch_wave_threshold_choices_agree(H, monkeypatch, wave_row):
    """Every wave threshold the autotuner may pick solves the instances to
    the same optimum (rows of >= 16 entries reduced by a wavefront, or never)."""
    qps = Q.batched_instances(count=10, seed=5, n_lo=40, n_hi=600)
    models = [build(H, qp) for qp in qps]
    monkeypatch.setenv("HPRQP_BATCH_WAVE_ROW", wave_row)
    b = H.Batch(models, H.Parameters(verbose=False, stoptol=1e-8))
    assert b.wave_row == int(wave_row)
    b.solve(50)
    sol = b.solution()
    on = 0
    for k, model in enumerate(models):
        plain = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-8))
        n = model.num_variables()
        assert sol["status"][k] == plain.status == "OPTIMAL"
        assert abs(sol["objective"][k] - plain.objective) <= 1e-7 * max(1.0, abs(plain.objective)), k
        assert np.max(np.abs(sol["x"][on:on + n] - plain.x)) <= 1e-5 * max(1.0, np.max(np.abs(plain.x))), k
        on += n


def test_batch_autotuned_wave_threshold_is_a_candidate(H, monkeypatch):
    monkeypatch.delenv("HPRQP_BATCH_WAVE_ROW", raising=False)
    qps = Q.batched_instances(count=16, seed=2, n_lo=50, n_hi=800)
    b = H.Batch([build(H, qp) for qp in qps], H.Parameters(verbose=False, stoptol=1e-6))
    assert b.wave_row in (16, 32, 64, 128, 1 << 20)
    assert b.solve(100) == 0
