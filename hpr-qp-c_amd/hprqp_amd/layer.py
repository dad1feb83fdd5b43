"""Differentiable QP layer on PyTorch-ROCm over the batched HIP solver.

    x = QPLayer.apply(Q, c, A, AL, AU, l, u)   # batched, FP64

solves, for every batch entry, min 1/2 x'Qx + c'x s.t. AL <= Ax <= AU,
l <= x <= u (equality rows: AL = AU) with the batched persistent HIP kernel,
and differentiates the converged KKT point implicitly: the active set is read
off the projections (nonzero y_bar / z_bar), and the active-set KKT system is
solved matrix-free by the HIP adjoint kernel (hprqp_hip_batch_adjoint). With
K [u; v] = [dL/dx; 0]:
    dL/dc = -u,  dL/dQ = -(u x' + x u')/2,  dL/dA = y u' - v_rows x',
    dL/dAL_i = v_i (row active at its lower bound), dL/dAU_i = v_i (upper),
    equality rows split v_i evenly between AL_i and AU_i (so d/db = v_i),
    dL/dl_j, dL/du_j likewise from the bound multipliers.

Tensors on a ROCm device stay there: the batch is created from their device
memory (hprqp_hip_batch_create_dense_device, ordered after torch's current
stream), the solution and the adjoint are written into torch tensors by
kernels on that stream, and no array crosses PCIe. CPU tensors go through
host arrays.
"""
from __future__ import annotations

import time
import warnings

import numpy as np
import torch

from . import Batch, Parameters, batch_adjoint_arrays, batch_adjoint_device, batch_adjoint_residuals


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


# Wall-clock split of the last forward / backward (seconds), for benchmarks.
last_timings: dict = {}

# The adjoint's MINRES stops when its residual estimate reaches ADJOINT_TOL
# (relative) or after ADJOINT_MAX_STEPS steps; the kernel then measures the
# true residual. A backward whose worst instance ends above ADJOINT_WARN
# warns (the gradients are as accurate as that residual allows); above
# ADJOINT_FAIL it raises, naming the instance.
ADJOINT_TOL = 1e-12
ADJOINT_MAX_STEPS = 20000
ADJOINT_WARN = 1e-10
ADJOINT_FAIL = 1e-6


def _check_adjoint(residuals: np.ndarray) -> None:
    worst = int(np.argmax(residuals)) if len(residuals) else 0
    r = float(residuals[worst]) if len(residuals) else 0.0
    last_timings["adjoint_max_residual"] = r
    if r > ADJOINT_FAIL:
        raise RuntimeError(f"QPLayer backward: the adjoint of instance {worst} did not converge (relative residual "
                           f"{r:.3e} after at most {ADJOINT_MAX_STEPS} steps); its gradients are not reliable")
    if r > ADJOINT_WARN:
        bad = int(np.count_nonzero(residuals > ADJOINT_WARN))
        warnings.warn(f"QPLayer backward: {bad} adjoint solve(s) ended above the relative residual {ADJOINT_WARN:g} "
                      f"(worst: instance {worst}, relative residual {r:.3e})", RuntimeWarning, stacklevel=3)


class QPLayer(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, c, A, AL, AU, l, u, stoptol: float = 1e-8, bodies_per_launch: int = 200):
        if Q.is_cuda:
            return _forward_device(ctx, (Q, c, A, AL, AU, l, u), stoptol, bodies_per_launch)
        ctx.on_device = False
        t0 = time.perf_counter()
        Qn, cn, An, ALn, AUn, ln, un = (_np(t) for t in (Q, c, A, AL, AU, l, u))
        B, n = cn.shape
        m = An.shape[1]
        t1 = time.perf_counter()
        try:
            # instances go to the device as dense arrays; build_from_QAbc's
            # preprocessing (Q symmetrised) runs there
            batch = Batch.from_dense(Qn, cn, An, ALn, AUn, ln, un, Parameters(verbose=False, stoptol=stoptol))
        except ValueError as e:
            raise ValueError("QPLayer: rows with no entries or free bounds are not supported "
                             f"(or invalid data): {e}") from None
        t2 = time.perf_counter()
        batch.solve(bodies_per_launch)
        t3 = time.perf_counter()
        sol = batch.solution()
        bad = [s_ for s_ in sol["status"] if s_ != "OPTIMAL"]
        if bad:
            raise RuntimeError(f"QPLayer forward: {len(bad)} instances not optimal ({bad[0]})")
        x = sol["x"].reshape(B, n)
        y = sol["y"].reshape(B, m)
        z = sol["z"].reshape(B, n)
        ctx.batch = batch
        ctx.sol = (x, y, z, ALn, AUn, ln, un)
        last_timings.clear()
        last_timings.update(to_host=t1 - t0, create=t2 - t1, solve=t3 - t2, device_solve=batch.seconds,
                            collect=time.perf_counter() - t3, iterations=int(sol["iter"].max()),
                            iterations_sum=int(sol["iter"].sum()))
        ctx.device = Q.device
        return torch.as_tensor(x, dtype=torch.float64, device=Q.device)

    @staticmethod
    def backward(ctx, gx):
        if ctx.on_device:
            return _backward_device(ctx, gx)
        x, y, z, ALn, AUn, ln, un = ctx.sol
        B, n = x.shape
        m = y.shape[1]
        g = _np(gx)
        t0 = time.perf_counter()
        u, vr, vb, _ = batch_adjoint_arrays(ctx.batch, g.reshape(-1), ADJOINT_TOL, ADJOINT_MAX_STEPS)
        last_timings["adjoint"] = time.perf_counter() - t0
        _check_adjoint(batch_adjoint_residuals(ctx.batch))
        # rank-one gradient terms as batched outer products (torch runs them
        # multi-threaded on the host, or on the device the inputs came from)
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=ctx.device)
        U, VR, VB = t(u.reshape(B, n)), t(vr.reshape(B, m)), t(vb.reshape(B, n))
        X, Y, Z = t(x), t(y), t(z)
        eq, feq = t(ALn == AUn).bool(), t(ln == un).bool()
        grads = _gradients(U, VR, VB, X, Y, Z, eq, feq)
        last_timings["backward_host"] = time.perf_counter() - t0
        return grads + (None, None)


def _gradients(U, VR, VB, X, Y, Z, eq, feq):
    """dL/d(Q, c, A, AL, AU, l, u) from the adjoint (U, VR, VB) at the
    solution (X, Y, Z); eq / feq mark equality rows / fixed variables."""
    P = torch.bmm(U[:, :, None], X[:, None, :])
    dQ = (P + P.transpose(1, 2)).mul_(-0.5)
    dA = torch.baddbmm(torch.bmm(Y[:, :, None], U[:, None, :]), VR[:, :, None], X[:, None, :], alpha=-1.0)
    dc = -U
    zero = torch.zeros((), dtype=torch.float64, device=U.device)
    dAL = torch.where(eq, 0.5 * VR, torch.where(Y > 0, VR, zero))
    dAU = torch.where(eq, 0.5 * VR, torch.where(Y < 0, VR, zero))
    dl = torch.where(feq, 0.5 * VB, torch.where(Z > 0, VB, zero))
    du = torch.where(feq, 0.5 * VB, torch.where(Z < 0, VB, zero))
    return dQ, dc, dA, dAL, dAU, dl, du


def _forward_device(ctx, inputs, stoptol, bodies_per_launch):
    t0 = time.perf_counter()
    dev = inputs[0].device
    if any(t.device != dev for t in inputs):
        raise ValueError("QPLayer: all inputs must be on one device")
    ts = [t.detach().to(torch.float64).contiguous() for t in inputs]
    B, n = ts[1].shape
    m = ts[2].shape[1]
    if ts[0].shape != (B, n, n) or ts[2].shape != (B, m, n) or any(t.shape != (B, m) for t in ts[3:5]) or any(
            t.shape != (B, n) for t in ts[5:]):
        raise ValueError("QPLayer: inconsistent shapes")
    stream = torch.cuda.current_stream(dev).cuda_stream
    t1 = time.perf_counter()
    try:
        index = dev.index if dev.index is not None else torch.cuda.current_device()
        batch = Batch.from_dense_device(B, n, m, [t.data_ptr() for t in ts], stream,
                                        Parameters(verbose=False, stoptol=stoptol, device_number=index))
    except ValueError as e:
        raise ValueError("QPLayer: rows with no entries or free bounds are not supported "
                         f"(or invalid data): {e}") from None
    t2 = time.perf_counter()
    batch.solve(bodies_per_launch)
    t3 = time.perf_counter()
    info = batch.status()
    bad = [s_ for s_ in info["status"] if s_ != "OPTIMAL"]
    if bad:
        raise RuntimeError(f"QPLayer forward: {len(bad)} instances not optimal ({bad[0]})")
    x = torch.empty((B, n), dtype=torch.float64, device=dev)
    y = torch.empty((B, m), dtype=torch.float64, device=dev)
    z = torch.empty((B, n), dtype=torch.float64, device=dev)
    batch.solution_device(x.data_ptr(), y.data_ptr(), z.data_ptr(), stream)
    ctx.on_device = True
    ctx.batch = batch
    ctx.sol = (y, z, ts[3] == ts[4], ts[5] == ts[6])
    ctx.save_for_backward(x)
    last_timings.clear()
    last_timings.update(to_host=t1 - t0, create=t2 - t1, solve=t3 - t2, device_solve=batch.seconds,
                        collect=time.perf_counter() - t3, iterations=int(info["iter"].max()),
                        iterations_sum=int(info["iter"].sum()))
    return x


def _backward_device(ctx, gx):
    (x,) = ctx.saved_tensors
    y, z, eq, feq = ctx.sol
    B, n = x.shape
    m = y.shape[1]
    t0 = time.perf_counter()
    g = gx.detach().to(torch.float64).contiguous()
    U = torch.empty((B, n), dtype=torch.float64, device=x.device)
    VB = torch.empty_like(U)
    VR = torch.empty((B, m), dtype=torch.float64, device=x.device)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    batch_adjoint_device(ctx.batch, g.data_ptr(), U.data_ptr(), VR.data_ptr(), VB.data_ptr(), stream,
                         ADJOINT_TOL, ADJOINT_MAX_STEPS)
    # the gradient products are queued behind the adjoint before its
    # convergence is read back (one wait for `count` residuals)
    grads = _gradients(U, VR, VB, x, y, z, eq, feq)
    last_timings["adjoint"] = time.perf_counter() - t0
    _check_adjoint(batch_adjoint_residuals(ctx.batch, stream))
    last_timings["backward_host"] = time.perf_counter() - t0
    return grads + (None, None)


def qp_layer(Q, c, A, AL, AU, l, u, stoptol: float = 1e-8):
    return QPLayer.apply(Q, c, A, AL, AU, l, u, stoptol)
