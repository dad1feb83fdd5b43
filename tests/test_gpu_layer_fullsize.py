"""The differentiable layer at its full size: 1,024 dense QPs (n = 256,
64 equality + 256 two-sided rows, Q = MM'/n + 0.1 I, A ~ N(0, 1/n), bounds by
the x0 construction) as ROCm tensors, forward to stoptol 1e-8 and backward
for an upstream gradient dL/dx ~ N(0, 1) (its own random stream). On 16 instances spread over the
batch:

* the forward x equals that of a batch of only those instances bit for bit
  (instances are independent), and agrees with a dense FP64 solve of the
  KKT system on the active set read off the duals (within 1e-6 relative);
* the gradients in Q, c, A, AL and AU agree with the dense FP64 KKT-solve
  gradient (1e-6 relative to the instance's largest gradient entry) and,
  through the active-set KKT solution map, with central finite differences
  (h = 1e-6)
  to 1e-5 relative (with a floor of 1e-3 of that largest entry, for the
  instances whose active rows pin x and leave only rounding noise in the
  Q and c gradients): c, Q and A entries, an equality row's b (AL = AU moved
  together), the lower bound of a row active at its lower bound, the upper
  bound of a row active at its upper bound, and both bounds of an inactive
  row (zero).
"""
import numpy as np
import pytest

import qp_libs as Q

pytestmark = pytest.mark.gpu

B, N, MEQ, MIN = 1024, 256, 64, 256
SAMPLE = list(range(0, B, B // 16))


@pytest.fixture(scope="module")
def run():
    import hprqp_amd as H
    if not H.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    import torch
    from hprqp_amd.layer import QPLayer
    arrs = Q.layer_instances(B)
    dev = torch.device("cuda:0")
    ts = [torch.tensor(a, dtype=torch.float64, device=dev, requires_grad=True) for a in arrs]
    # upstream gradient from its own stream (the generator's seed-0 stream
    # would make instance 0's g a multiple of one of its own rows of A)
    g = np.random.default_rng(12345).normal(size=(B, N))
    print("layer_1024 forward", flush=True)
    x = QPLayer.apply(*ts, 1e-8)
    print("layer_1024 backward", flush=True)
    (x * torch.tensor(g, device=dev)).sum().backward()
    torch.cuda.synchronize()
    grads = [t.grad.cpu().numpy() for t in ts]
    # duals of the sampled instances from a batch holding only them
    sub = [a[SAMPLE] for a in arrs]
    sb = H.Batch.from_dense(*sub, H.Parameters(verbose=False, stoptol=1e-8))
    sb.solve(200)
    sol = sb.solution()
    m = MEQ + MIN
    ys = sol["y"].reshape(len(SAMPLE), m)
    zs = sol["z"].reshape(len(SAMPLE), N)
    xs = sol["x"].reshape(len(SAMPLE), N)
    return arrs, x.detach().cpu().numpy(), grads, g, xs, ys, zs


def test_layer_1024_forward_matches_subset_and_dense_kkt(run):
    arrs, x, grads, g, xs, ys, zs = run
    worst = 0.0
    for j, b in enumerate(SAMPLE):
        np.testing.assert_array_equal(x[b], xs[j], err_msg=f"instance {b}")
        Qd, c, A, AL, AU, l, u = (a[b] for a in arrs)
        xk = Q.kkt_active_solution(Qd, c, A, AL, AU, l, u, ys[j], zs[j])[0]
        err = np.max(np.abs(x[b] - xk)) / max(1.0, np.max(np.abs(xk)))
        worst = max(worst, err)
        assert err <= 1e-6, (b, err)
    print(f"forward vs dense KKT: worst {worst:.3e}", flush=True)


def _dense_grads(Qd, c, A, AL, AU, l, u, x, y, z, g):
    """Gradients from a dense FP64 solve of the active-set KKT system
    K [u; v] = [g; 0] at the solution (x, y, z)."""
    _, K, rows, vars_ = Q.kkt_active_solution(Qd, c, A, AL, AU, l, u, y, z)
    n = len(c)
    sol = np.linalg.solve(K, np.concatenate([g, np.zeros(K.shape[0] - n)]))
    uu = sol[:n]
    v = np.zeros(len(AL))
    v[rows] = sol[n:n + len(rows)]
    eq = AL == AU
    dAL = np.where(eq, 0.5 * v, np.where(y > 0, v, 0.0))
    dAU = np.where(eq, 0.5 * v, np.where(y < 0, v, 0.0))
    return {"Q": -0.5 * (np.outer(uu, x) + np.outer(x, uu)), "c": -uu, "A": np.outer(y, uu) - np.outer(v, x),
            "AL": dAL, "AU": dAU}


def test_layer_1024_gradients_match_dense_kkt_and_fd(run):
    arrs, x, grads, g, xs, ys, zs = run
    names = ["Q", "c", "A", "AL", "AU", "l", "u"]
    h = 1e-6
    checked = {"lower": 0, "upper": 0, "inactive": 0}
    for j, b in enumerate(SAMPLE):
        Qd, c, A, AL, AU, l, u = (a[b].copy() for a in arrs)
        y, z = ys[j], zs[j]
        ref = _dense_grads(Qd, c, A, AL, AU, l, u, x[b], y, z, g[b])
        got = {nm: grads[k][b] for k, nm in enumerate(names)}
        # one scale per instance: where the active rows pin x (as many
        # independent active rows as variables) u vanishes, and the gradients
        # in Q and c are rounding noise next to those in A and the bounds
        G = max(float(np.max(np.abs(w))) for w in ref.values())
        print(f"instance {b}: {int(np.sum(y != 0))} active rows of {MEQ + MIN}, max|dQ| "
              f"{float(np.max(np.abs(ref['Q']))):.2e}, max gradient {G:.2e}", flush=True)
        for nm, want in ref.items():
            assert np.max(np.abs(got[nm] - want)) <= 1e-6 * G, (b, nm, float(np.max(np.abs(got[nm] - want))), G)
        assert not np.any(got["l"]) and not np.any(got["u"])  # free variables

        def fd(which, idx, both=False):
            arr_p = {k: v.copy() for k, v in zip(names, (Qd, c, A, AL, AU, l, u))}
            arr_m = {k: v.copy() for k, v in zip(names, (Qd, c, A, AL, AU, l, u))}
            for w in (("AL", "AU") if both else (which,)):
                arr_p[w][idx] += h
                arr_m[w][idx] -= h
            xp = Q.kkt_active_solution(*(arr_p[k] for k in names), y, z)[0]
            xm = Q.kkt_active_solution(*(arr_m[k] for k in names), y, z)[0]
            return float(g[b] @ (xp - xm) / (2 * h))

        def close_fd(which, idx, val, want):
            scale = max(abs(want), 1e-3 * G)
            assert abs(val - want) <= 1e-5 * scale, (b, which, idx, val, want)

        rng = np.random.default_rng(b)
        jc = int(rng.integers(N))
        close_fd("c", jc, got["c"][jc], fd("c", jc))
        r_, s_ = int(rng.integers(N)), int(rng.integers(N))
        close_fd("Q", (r_, s_), got["Q"][r_, s_], fd("Q", (r_, s_)))
        i2, s2 = int(rng.integers(MEQ + MIN)), int(rng.integers(N))
        close_fd("A", (i2, s2), got["A"][i2, s2], fd("A", (i2, s2)))
        ie = int(rng.integers(MEQ))  # equality row: d/db = dAL + dAU
        want = fd("AL", ie, both=True)
        assert abs(got["AL"][ie] + got["AU"][ie] - want) <= 1e-5 * max(abs(want), 1e-3 * G)
        ineq = np.arange(MEQ, MEQ + MIN)
        for kind, mask in (("lower", y[ineq] > 0), ("upper", y[ineq] < 0), ("inactive", y[ineq] == 0)):
            rows = ineq[mask]
            if not len(rows):
                continue
            i = int(rows[rng.integers(len(rows))])
            for which in ("AL", "AU"):
                want = fd(which, i)
                if kind == "inactive" or (kind == "lower") != (which == "AL"):
                    assert want == 0.0 and got[which][i] == 0.0, (b, kind, which, i)
                else:
                    close_fd(which, i, got[which][i], want)
            checked[kind] += 1
    print("inequality rows checked:", checked, flush=True)
    assert min(checked.values()) >= 8, checked
