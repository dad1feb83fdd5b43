"""Differentiable QP layer: the HIP adjoint (hprqp_hip_batch_adjoint) against
(1) a dense FP64 KKT solve of the same active-set system and (2) central
finite differences (h = 1e-6) of the active-set KKT solution map, and the
forward solution against that dense solve."""
import math

import numpy as np
import pytest

import qp_libs as Q

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    import torch
    from hprqp_amd.layer import QPLayer
    return hprqp_amd, torch, QPLayer


def dense_grads(Qm, c, A, AL, AU, l, u, x, y, z, g):
    xs, K, rows, vars_ = Q.kkt_active_solution(Qm, c, A, AL, AU, l, u, y, z)
    n = len(c)
    sol = np.linalg.solve(K, np.concatenate([g, np.zeros(K.shape[0] - n)]))
    uu = sol[:n]
    v = np.zeros(len(AL))
    v[rows] = sol[n:n + len(rows)]
    dQ = -0.5 * (np.outer(uu, x) + np.outer(x, uu))
    dA = np.outer(y, uu) - np.outer(v, x)
    return xs, dict(dc=-uu, dQ=dQ, dA=dA, v=v)


def test_layer_gradients_match_dense_kkt_and_fd(mods):
    H, torch, QPLayer = mods
    B, n, meq, min_ = 4, 24, 6, 20
    Qd, c, A, AL, AU, l, u = Q.layer_instances(B, n=n, m_eq=meq, m_in=min_, seed=3)
    T = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)
    tQ, tc, tA, tAL, tAU, tl, tu = (T(a) for a in (Qd, c, A, AL, AU, l, u))
    x = QPLayer.apply(tQ, tc, tA, tAL, tAU, tl, tu, 1e-10)
    g = np.random.default_rng(0).normal(size=(B, n))
    (x * torch.tensor(g)).sum().backward()
    for b in range(B):
        xs_h = x[b].detach().numpy()
        # duals for the active set of the dense reference: a plain solve of the instance
        model = H.build_from_qabc(Qd[b], c[b], A[b], AL[b], AU[b], l[b], u[b])
        pr = H.optimize(model, H.Parameters(verbose=False, stoptol=1e-10))
        assert pr.status == "OPTIMAL"
        xs, ref = dense_grads(Qd[b], c[b], A[b], AL[b], AU[b], l[b], u[b], pr.x, pr.y, pr.z, g[b])
        assert np.max(np.abs(xs_h - xs)) <= 1e-6 * max(1.0, np.max(np.abs(xs)))
        for name, got in (("dc", tc.grad[b]), ("dQ", tQ.grad[b]), ("dA", tA.grad[b])):
            want = ref[name]
            assert np.max(np.abs(got.numpy() - want)) <= 1e-6 * max(1.0, np.max(np.abs(want))), (b, name)
        # central differences of the active-set KKT solution map, h = 1e-6
        h = 1e-6
        rng = np.random.default_rng(b)
        for _ in range(4):
            j = int(rng.integers(n))
            cp, cm = c[b].copy(), c[b].copy()
            cp[j] += h
            cm[j] -= h
            xp = Q.kkt_active_solution(Qd[b], cp, A[b], AL[b], AU[b], l[b], u[b], pr.y, pr.z)[0]
            xm = Q.kkt_active_solution(Qd[b], cm, A[b], AL[b], AU[b], l[b], u[b], pr.y, pr.z)[0]
            fd = g[b] @ (xp - xm) / (2 * h)
            assert abs(tc.grad[b, j].item() - fd) <= 1e-5 * max(1.0, abs(fd)), ("dc", j)
            i = int(rng.integers(meq))  # equality row: d/db = dAL + dAU
            bp = [AL[b].copy(), AU[b].copy()]
            bm = [AL[b].copy(), AU[b].copy()]
            for arr in bp:
                arr[i] += h
            for arr in bm:
                arr[i] -= h
            xp = Q.kkt_active_solution(Qd[b], c[b], A[b], bp[0], bp[1], l[b], u[b], pr.y, pr.z)[0]
            xm = Q.kkt_active_solution(Qd[b], c[b], A[b], bm[0], bm[1], l[b], u[b], pr.y, pr.z)[0]
            fd = g[b] @ (xp - xm) / (2 * h)
            got = tAL.grad[b, i].item() + tAU.grad[b, i].item()
            assert abs(got - fd) <= 1e-5 * max(1.0, abs(fd)), ("db", i)
            r_, s_ = int(rng.integers(n)), int(rng.integers(n))
            Qp, Qm_ = Qd[b].copy(), Qd[b].copy()
            Qp[r_, s_] += h
            Qm_[r_, s_] -= h
            xp = Q.kkt_active_solution(Qp, c[b], A[b], AL[b], AU[b], l[b], u[b], pr.y, pr.z)[0]
            xm = Q.kkt_active_solution(Qm_, c[b], A[b], AL[b], AU[b], l[b], u[b], pr.y, pr.z)[0]
            fd = g[b] @ (xp - xm) / (2 * h)
            assert abs(tQ.grad[b, r_, s_].item() - fd) <= 1e-5 * max(1.0, abs(fd)), ("dQ", r_, s_)
            i2, s2 = int(rng.integers(meq + min_)), int(rng.integers(n))
            Ap, Am = A[b].copy(), A[b].copy()
            Ap[i2, s2] += h
            Am[i2, s2] -= h
            xp = Q.kkt_active_solution(Qd[b], c[b], Ap, AL[b], AU[b], l[b], u[b], pr.y, pr.z)[0]
            xm = Q.kkt_active_solution(Qd[b], c[b], Am, AL[b], AU[b], l[b], u[b], pr.y, pr.z)[0]
            fd = g[b] @ (xp - xm) / (2 * h)
            assert abs(tA.grad[b, i2, s2].item() - fd) <= 1e-5 * max(1.0, abs(fd)), ("dA", i2, s2)


def test_layer_solver_fd_on_c(mods):
    """Finite differences through the HIP solver itself (stoptol 1e-11)."""
    H, torch, QPLayer = mods
    Qd, c, A, AL, AU, l, u = Q.layer_instances(1, n=16, m_eq=4, m_in=12, seed=9)
    tc = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    args = [torch.tensor(a, dtype=torch.float64) for a in (Qd, A, AL, AU, l, u)]
    x = QPLayer.apply(args[0], tc, args[1], args[2], args[3], args[4], args[5], 1e-11)
    g = np.random.default_rng(1).normal(size=(1, 16))
    (x * torch.tensor(g)).sum().backward()
    h = 1e-6
    for j in range(0, 16, 5):
        cp, cm = c.copy(), c.copy()
        cp[0, j] += h
        cm[0, j] -= h
        xp = QPLayer.apply(args[0], torch.tensor(cp), *args[1:], 1e-11).numpy()
        xm = QPLayer.apply(args[0], torch.tensor(cm), *args[1:], 1e-11).numpy()
        fd = float(g[0] @ (xp[0] - xm[0]) / (2 * h))
        assert abs(tc.grad[0, j].item() - fd) <= 1e-3 * max(1.0, abs(fd)), (j, tc.grad[0, j].item(), fd)


def test_device_batch_equals_host_batch(mods):
    """hprqp_hip_batch_create_dense_device on device tensors builds the same
    batch as the host-array entry point: solutions (device writer vs host
    copy) and adjoints (device buffers vs host arrays) bit-identical."""
    H, torch, _ = mods
    B, n, meq, min_ = 6, 40, 8, 30
    arrs = Q.layer_instances(B, n=n, m_eq=meq, m_in=min_, seed=5)
    m = meq + min_
    print("device batch", B, n, m, flush=True)
    hb = H.Batch.from_dense(*arrs, H.Parameters(verbose=False, stoptol=1e-9))
    hb.solve()
    hs = hb.solution()
    dev = torch.device("cuda:0")
    ts = [torch.tensor(a, dtype=torch.float64, device=dev) for a in arrs]
    stream = torch.cuda.current_stream(dev).cuda_stream
    db = H.Batch.from_dense_device(B, n, m, [t.data_ptr() for t in ts], stream,
                                   H.Parameters(verbose=False, stoptol=1e-9))
    db.solve()
    info = db.status()
    assert info["status"] == hs["status"] and np.array_equal(info["iter"], hs["iter"])
    assert np.array_equal(info["objective"], hs["objective"])
    x, y, z = (torch.full((B * k,), np.nan, dtype=torch.float64, device=dev) for k in (n, m, n))
    db.solution_device(x.data_ptr(), y.data_ptr(), z.data_ptr(), stream)
    torch.cuda.synchronize()
    for name, got in (("x", x), ("y", y), ("z", z)):
        np.testing.assert_array_equal(got.cpu().numpy(), hs[name], err_msg=name)
    g = np.random.default_rng(1).normal(size=B * n)
    u_h, vr_h, vb_h, steps_h = H.batch_adjoint_arrays(hb, g)
    tg = torch.tensor(g, device=dev)
    U, VB, VR = (torch.full((B * k,), np.nan, dtype=torch.float64, device=dev) for k in (n, n, m))
    steps = H.batch_adjoint_device(db, tg.data_ptr(), U.data_ptr(), VR.data_ptr(), VB.data_ptr(), stream,
                                   want_steps=True)
    torch.cuda.synchronize()
    assert np.array_equal(steps, steps_h)
    for name, got, want in (("u", U, u_h), ("v_rows", VR, vr_h), ("v_bounds", VB, vb_h)):
        np.testing.assert_array_equal(got.cpu().numpy(), want, err_msg=name)


def test_device_create_rejects_bad_input(mods):
    """Validation on the device: a NaN in Q, a row of A without entries and
    an inverted bound are rejected like the host entry point rejects them;
    host memory is refused as input."""
    H, torch, _ = mods
    B, n, meq, min_ = 3, 16, 4, 8
    arrs = [np.array(a) for a in Q.layer_instances(B, n=n, m_eq=meq, m_in=min_, seed=2)]
    m = meq + min_
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    params = H.Parameters(verbose=False)

    def create(a):
        ts = [torch.tensor(v, dtype=torch.float64, device=dev) for v in a]
        return H.Batch.from_dense_device(B, n, m, [t.data_ptr() for t in ts], stream, params)
    assert len(create(arrs)) == B
    for k, (idx, val) in enumerate([((0, (1, 2, 3)), np.nan), ((2, (2, 5)), "zero_row"), ((4, (1, 3)), "inverted")]):
        bad = [a.copy() for a in arrs]
        which, pos = idx
        if val == "zero_row":
            bad[which][pos[0], pos[1], :] = 0.0
        elif val == "inverted":
            bad[3][pos] = bad[4][pos] + 1.0
        else:
            bad[which][pos] = val
        print("reject case", k, flush=True)
        with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
            create(bad)
        with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
            H.Batch.from_dense(*bad, params)
    host = [np.ascontiguousarray(a, np.float64) for a in arrs]
    with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
        H.Batch.from_dense_device(B, n, m, [a.ctypes.data for a in host], stream, params)


def test_layer_on_device_tensors_matches_host_tensors(mods):
    """QPLayer on ROCm tensors (zero-copy path) against the same layer on CPU
    tensors: forward bit-identical, gradients to 1e-12 (the batched torch
    products run on different devices)."""
    H, torch, QPLayer = mods
    B, n, meq, min_ = 5, 32, 6, 24
    arrs = Q.layer_instances(B, n=n, m_eq=meq, m_in=min_, seed=9)
    g = np.random.default_rng(2).normal(size=(B, n))
    out = {}
    for where in ("cpu", "cuda:0"):
        print("layer on", where, flush=True)
        ts = [torch.tensor(a, dtype=torch.float64, device=where, requires_grad=True) for a in arrs]
        x = QPLayer.apply(*ts, 1e-9)
        assert x.device.type == torch.device(where).type
        (x * torch.tensor(g, device=where)).sum().backward()
        out[where] = [x.detach().cpu().numpy()] + [t.grad.cpu().numpy() for t in ts]
    np.testing.assert_array_equal(out["cuda:0"][0], out["cpu"][0])
    for k, name in enumerate(("Q", "c", "A", "AL", "AU", "l", "u"), start=1):
        want, got = out["cpu"][k], out["cuda:0"][k]
        assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want))), name


def test_device_outputs_complete_when_batch_destroyed(mods):
    """The device-buffer solution and adjoint are queued on the caller's
    stream; destroying the batch right after waits for them, so the outputs
    are whole."""
    H, torch, _ = mods
    B, n, meq, min_ = 4, 24, 4, 16
    arrs = Q.layer_instances(B, n=n, m_eq=meq, m_in=min_, seed=11)
    m = meq + min_
    hb = H.Batch.from_dense(*arrs, H.Parameters(verbose=False, stoptol=1e-9))
    hb.solve()
    want = hb.solution()
    g = np.random.default_rng(3).normal(size=B * n)
    u_h = H.batch_adjoint_arrays(hb, g)[0]
    dev = torch.device("cuda:0")
    ts = [torch.tensor(a, dtype=torch.float64, device=dev) for a in arrs]
    stream = torch.cuda.current_stream(dev).cuda_stream
    db = H.Batch.from_dense_device(B, n, m, [t.data_ptr() for t in ts], stream, H.Parameters(verbose=False, stoptol=1e-9))
    db.solve()
    x = torch.empty(B * n, dtype=torch.float64, device=dev)
    tg = torch.tensor(g, device=dev)
    U, VB = torch.empty_like(x), torch.empty_like(x)
    VR = torch.empty(B * m, dtype=torch.float64, device=dev)
    db.solution_device(x.data_ptr(), 0, 0, stream)
    H.batch_adjoint_device(db, tg.data_ptr(), U.data_ptr(), VR.data_ptr(), VB.data_ptr(), stream)
    del db  # no synchronisation by the caller before the batch goes
    torch.cuda.synchronize()
    np.testing.assert_array_equal(x.cpu().numpy(), want["x"])
    np.testing.assert_array_equal(U.cpu().numpy(), u_h)


def _boxed_instances(count, n, m_eq, m_in, seed):
    """Layer instances with finite variable bounds around the generating
    point (x0 -/+ U(0, 0.3)), so that the optimum has variables at their
    lower bound, at their upper bound and strictly inside, plus one fixed
    variable (l = u) per instance."""
    Qd, c, A, AL, AU, l, u = Q.layer_instances(count, n=n, m_eq=m_eq, m_in=m_in, seed=seed)
    rng = np.random.default_rng(seed + 100)
    c = c * 4.0  # push the optimum onto the bounds
    l, u = np.empty((count, n)), np.empty((count, n))
    for k in range(count):
        x0 = rng.uniform(-1.0, 1.0, size=n)
        ax = A[k] @ x0
        AL[k, :m_eq] = AU[k, :m_eq] = ax[:m_eq]
        AL[k, m_eq:] = ax[m_eq:] - rng.uniform(0.05, 0.5, size=m_in)
        AU[k, m_eq:] = ax[m_eq:] + rng.uniform(0.05, 0.5, size=m_in)
        l[k] = x0 - rng.uniform(0.02, 0.3, size=n)
        u[k] = x0 + rng.uniform(0.02, 0.3, size=n)
        l[k, 0] = u[k, 0] = x0[0]
    return Qd, c, A, AL, AU, l, u


def test_layer_bound_and_row_gradients_match_fd(mods):
    """dL/dl, dL/du and dL/dAL, dL/dAU through QPLayer against central
    differences (h = 1e-6) of the active-set KKT solution map, 1e-5
    relative: for a variable at its lower bound only l moves x, at its upper
    bound only u, strictly inside neither (both gradients exactly 0); a
    fixed variable (l = u) moves with l and u together; likewise for rows at
    their lower / upper bound, inactive rows and equality rows."""
    H, torch, QPLayer = mods
    count, n, meq, min_ = 4, 40, 6, 24
    arrs = _boxed_instances(count, n, meq, min_, seed=17)
    names = ["Q", "c", "A", "AL", "AU", "l", "u"]
    T = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)
    ts = [T(a) for a in arrs]
    x = QPLayer.apply(*ts, 1e-10)
    g = np.random.default_rng(4).normal(size=(count, n))
    (x * torch.tensor(g)).sum().backward()
    grads = {nm: t.grad.numpy() for nm, t in zip(names, ts)}
    b = H.Batch.from_dense(*arrs, H.Parameters(verbose=False, stoptol=1e-10))
    b.solve()
    sol = b.solution()
    ys, zs = sol["y"].reshape(count, meq + min_), sol["z"].reshape(count, n)
    np.testing.assert_array_equal(sol["x"].reshape(count, n), x.detach().numpy())
    h = 1e-6
    seen = {"l": 0, "u": 0, "free": 0, "AL": 0, "AU": 0, "inactive": 0}
    for k in range(count):
        base = {nm: a[k].copy() for nm, a in zip(names, arrs)}
        y, z = ys[k], zs[k]

        def fd(moves):
            p = {nm: v.copy() for nm, v in base.items()}
            m_ = {nm: v.copy() for nm, v in base.items()}
            for nm, idx in moves:
                p[nm][idx] += h
                m_[nm][idx] -= h
            xp = Q.kkt_active_solution(*(p[nm] for nm in names), y, z)[0]
            xm = Q.kkt_active_solution(*(m_[nm] for nm in names), y, z)[0]
            return float(g[k] @ (xp - xm) / (2 * h))

        def check(got, want, nm):
            scale = max(abs(want), 1e-2 * float(np.max(np.abs(grads[nm][k]))), 1e-12)
            assert abs(got - want) <= 1e-5 * scale, (k, nm, got, want)
        # the fixed variable 0: l and u moved together
        check(grads["l"][k][0] + grads["u"][k][0], fd([("l", 0), ("u", 0)]), "l")
        for j in range(1, n):
            kind = "l" if z[j] > 0 else ("u" if z[j] < 0 else "free")
            seen[kind] += 1
            for nm in ("l", "u"):
                want = fd([(nm, j)])
                if kind == nm:
                    check(grads[nm][k][j], want, nm)
                else:
                    assert want == 0.0 and grads[nm][k][j] == 0.0, (k, j, kind, nm)
        for i in range(meq):
            check(grads["AL"][k][i] + grads["AU"][k][i], fd([("AL", i), ("AU", i)]), "AL")
        for i in range(meq, meq + min_):
            kind = "AL" if y[i] > 0 else ("AU" if y[i] < 0 else "inactive")
            seen[kind] += 1
            for nm in ("AL", "AU"):
                want = fd([(nm, i)])
                if kind == nm:
                    check(grads[nm][k][i], want, nm)
                else:
                    assert want == 0.0 and grads[nm][k][i] == 0.0, (k, i, kind, nm)
    print("active sets seen:", seen, flush=True)
    assert all(v >= 2 for v in seen.values()), seen
