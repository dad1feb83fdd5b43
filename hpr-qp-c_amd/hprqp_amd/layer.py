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
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import Batch, Parameters, batch_adjoint, build_from_qabc


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


# Wall-clock split of the last forward / backward (seconds), for benchmarks.
last_timings: dict = {}


def _build_all(Qn, cn, An, ALn, AUn, ln, un):
    # build_from_QAbc is a ctypes call, which releases the GIL: build the
    # per-instance models (host symmetrisation and row checks) on threads.
    B = cn.shape[0]
    workers = max(1, min(16, os.cpu_count() or 1, B))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda b: build_from_qabc(Qn[b], cn[b], An[b], ALn[b], AUn[b], ln[b], un[b]), range(B)))


class QPLayer(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, c, A, AL, AU, l, u, stoptol: float = 1e-8, bodies_per_launch: int = 200):
        t0 = time.perf_counter()
        Qn, cn, An, ALn, AUn, ln, un = (_np(t) for t in (Q, c, A, AL, AU, l, u))
        B, n = cn.shape
        m = An.shape[1]
        models = _build_all(Qn, cn, An, ALn, AUn, ln, un)
        for md in models:
            if md.num_constraints() != m:
                raise ValueError("QPLayer: rows with no entries or free bounds are not supported")
        t1 = time.perf_counter()
        batch = Batch(models, Parameters(verbose=False, stoptol=stoptol))
        t2 = time.perf_counter()
        batch.solve(bodies_per_launch)
        t3 = time.perf_counter()
        res = [batch.result(b) for b in range(B)]
        bad = [r.status for r in res if r.status != "OPTIMAL"]
        if bad:
            raise RuntimeError(f"QPLayer forward: {len(bad)} instances not optimal ({bad[0]})")
        x = np.stack([r.x for r in res])
        y = np.stack([r.y for r in res]) if m else np.zeros((B, 0))
        z = np.stack([r.z for r in res])
        ctx.batch = batch
        ctx.sol = (x, y, z, ALn, AUn, ln, un)
        last_timings.clear()
        last_timings.update(build=t1 - t0, upload=t2 - t1, solve=t3 - t2, device_solve=batch.seconds,
                            collect=time.perf_counter() - t3, iterations=max(r.iter for r in res))
        ctx.device = Q.device
        return torch.as_tensor(x, dtype=torch.float64, device=Q.device)

    @staticmethod
    def backward(ctx, gx):
        x, y, z, ALn, AUn, ln, un = ctx.sol
        B, n = x.shape
        m = y.shape[1]
        g = _np(gx)
        t0 = time.perf_counter()
        adj = batch_adjoint(ctx.batch, list(g))
        last_timings["adjoint"] = time.perf_counter() - t0
        U = np.stack([a[0] for a in adj])
        VR = np.stack([a[1] for a in adj]) if m else np.zeros((B, 0))
        VB = np.stack([a[2] for a in adj])
        dc = -U
        dQ = -0.5 * (U[:, :, None] * x[:, None, :] + x[:, :, None] * U[:, None, :])
        dA = y[:, :, None] * U[:, None, :] - VR[:, :, None] * x[:, None, :]
        eq = ALn == AUn
        dAL = np.where(eq, 0.5 * VR, np.where(y > 0, VR, 0.0))
        dAU = np.where(eq, 0.5 * VR, np.where(y < 0, VR, 0.0))
        feq = ln == un
        dl = np.where(feq, 0.5 * VB, np.where(z > 0, VB, 0.0))
        du = np.where(feq, 0.5 * VB, np.where(z < 0, VB, 0.0))
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=ctx.device)
        return t(dQ), t(dc), t(dA), t(dAL), t(dAU), t(dl), t(du), None, None


def qp_layer(Q, c, A, AL, AU, l, u, stoptol: float = 1e-8):
    return QPLayer.apply(Q, c, A, AL, AU, l, u, stoptol)
