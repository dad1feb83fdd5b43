"""Benchmark: time to KKT < 1e-8 of one HPR-QP solve on one MI355X.

Workload (default, the plain solver at scale): a seeded random sparse convex
QP with n = 1,000,000 variables, m = 500,000 rows (300k equality, 200k
two-sided), A at 20 nnz/row with 1% of rows (5,000) at 2,000 nnz, Q = B'DB +
1e-2 I with B 200,000 x n at 8 nnz/row, box [-10, 10], FP64, Ruiz scaling on,
stoptol 1e-8, check_iter 100, print_frequency -1, no polishing, no Gram.

A "step" is one full solve from a fresh device session (upload, device
transpose, scaling, power iteration) to stoptol; the metric value is the
solve-loop time to tolerance (the reference's results.time), lower is better.
--warmup W runs W short solves (max_iter 50) first.

Besides the metric line it reports:
  roofline      the dominant iteration kernel's algorithmic bytes per launch /
                its mean launch time (HIP events on the library's own stream)
  cpu_baseline  the oracle (plain-C restatement of the reference CPU path, one
                core) timed on a bounded sample of the same workload: the
                first loop bodies after setup, scaled to the GPU solve's
                iteration count.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "hpr-qp-c_amd"))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0

SOC_WORKLOAD = dict(seed=0, n_box=20_000, cones=20_000, cone_dim=10, m=100_000, a_row_nnz=40, box=5.0)

WORKLOADS = {
    "plain_n1M": dict(seed=0, n=1_000_000, m_eq=300_000, m_in=200_000, a_row_nnz=20, heavy_rows=5_000,
                      heavy_nnz=2_000, b_rows=200_000, b_row_nnz=8, box=10.0),
    "plain_n2k": dict(seed=0, n=2_000, m_eq=600, m_in=400, a_row_nnz=10, heavy_rows=0, heavy_nnz=0,
                      b_rows=500, b_row_nnz=5, box=10.0),
}


def banded_variant(qp):
    """Same row lengths as `qp` (and the same Q row density), but each row's
    columns form one contiguous run near the diagonal: the SpMV kernels then
    gather almost nothing from L2/MALL. Profiling only (--profile-only), to
    separate the cost of random gathers from the cost of streaming."""
    n, m = qp.n, qp.m
    lens = np.diff(qp.a_ptr.astype(np.int64))
    starts = np.repeat((np.arange(m) * (n // max(m, 1))).astype(np.int64), lens)
    within = np.arange(len(qp.a_ind), dtype=np.int64) - np.repeat(qp.a_ptr[:-1].astype(np.int64), lens)
    cols = (starts + within) % n
    order = np.lexsort((cols, np.repeat(np.arange(m), lens)))
    qp.a_ind = cols[order].astype(np.uint64)
    qp.a_val = qp.a_val[order]
    half = 6
    rows = np.repeat(np.arange(n), 2 * half + 1)
    cols = rows + np.tile(np.arange(-half, half + 1), n)
    ok = (cols >= 0) & (cols < n)
    rows, cols = rows[ok], cols[ok]
    qp.q_ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.uint64)
    qp.q_ind = cols.astype(np.uint64)
    qp.q_val = np.where(rows == cols, 20.0, 1.0 / (1.0 + np.abs(rows - cols)))
    return qp


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_model(H, qp):
    return H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                             (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u)


def profile_kernels(H, model, iters):
    s = H.Session(model, H.Parameters(verbose=False))
    cap = 16
    names = C.create_string_buffer(32 * cap)
    ms = np.zeros(cap)
    by = np.zeros(cap)
    cnt = np.zeros(cap, np.uint64)
    lib = H.lib()
    lib.hprqp_hip_session_profile.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_char_p,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_size_t)]
    k = lib.hprqp_hip_session_profile(s.h, iters, cap, names, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                      by.ctypes.data_as(C.POINTER(C.c_double)),
                                      cnt.ctypes.data_as(C.POINTER(C.c_size_t)))
    out = []
    for t in range(max(k, 0)):
        nm = names.raw[32 * t:32 * (t + 1)].split(b"\0")[0].decode()
        if cnt[t]:
            out.append(dict(kernel=nm, mean_ms=float(ms[t]), bytes=float(by[t]), launches=int(cnt[t]),
                            gbs=float(by[t]) / (ms[t] * 1e-3) / 1e9))
    return out


def cpu_baseline(qp, gpu_iters, lam_a, lam_q, sample_bodies):
    import qp_libs as Q
    if sample_bodies >= 10 ** 9:  # small workload: the whole solve on one core
        if Q.ref_available():
            # the reference itself, built from its own sources (oracle/_ref),
            # through its binding API: solve_qp --use-gpu 0 equivalent
            t0 = time.time()
            r = Q.binding_solve(Q.ref_lib(), qp, use_gpu=0)
            dt = time.time() - t0
            return dict(value=r["time"], unit="s (solve loop to 1e-8)", cores=1, kind="reference",
                        sample=f"reference CPU path (oracle/_ref/libhprqp_ref.so), full solve ({r['iter']} "
                               f"iterations, {r['status']}); setup+solve {dt:.3f} s")
        t0 = time.time()
        o = Q.OracleSession(qp)
        r = o.solve()[0]
        dt = time.time() - t0
        return dict(value=r.time, unit="s (solve loop to 1e-8)", cores=1, kind="port",
                    sample=f"oracle, full solve ({r.iter} iterations, {Q.STATUS[r.status]}); "
                           f"setup+solve {dt:.3f} s")
    t0 = time.time()
    o = Q.OracleSession(qp, lambda_A_given=lam_a, lambda_Q_given=lam_q)
    setup = time.time() - t0
    o.advance(1)                      # body 0: iteration-0 residuals + first step
    t1 = time.time()
    o.advance(sample_bodies)
    dt = time.time() - t1
    per_iter = dt / sample_bodies
    return dict(value=per_iter * gpu_iters, unit="s (est. time to 1e-8)", cores=1, kind="port",
                sample=f"oracle, loop bodies 1..{sample_bodies} of this workload after setup "
                       f"({setup:.1f} s, power iteration skipped with the GPU's lambda), "
                       f"{per_iter * 1e3:.1f} ms/iteration x {gpu_iters} iterations")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain_n1M",
                    choices=sorted(WORKLOADS) + ["plain_n1M_banded", "soc_n220k", "batched_4096", "layer_1024"])
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--stoptol", type=float, default=1e-8)
    ap.add_argument("--profile-iters", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="set up the workload, run --profile-iters timed iterations and a 256 MiB device copy "
                         "(the reference kernel for counter calibration); no solve, no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is design-only.", file=sys.stderr)
        sys.exit(2)

    import hprqp_amd as H
    import qp_libs as Q
    if not H.gpu_backend_available():
        raise SystemExit("no gfx950 device visible")
    log("device:", H.device_name())
    if args.workload == "batched_4096":
        return bench_batched(H, Q, args)
    if args.workload == "layer_1024":
        return bench_layer(H, Q, args)
    if args.workload == "soc_n220k":
        qp = Q.soc_qp(**SOC_WORKLOAD)
        log(f"generated soc_n220k: n={qp.n} m={qp.m} nnzA={len(qp.a_val)} cones={qp.num_var_cones}")
        return bench_single(H, Q, args, qp, "soc_n220k", soc=True)
    banded = args.workload.endswith("_banded")
    if banded and not args.profile_only:
        raise SystemExit("the banded variant is a profiling aid: use it with --profile-only")
    kw = WORKLOADS[args.workload.replace("_banded", "")]
    t = time.time()
    qp = Q.random_qp(**kw)
    if banded:
        qp = banded_variant(qp)
    log(f"generated {args.workload}: n={qp.n} m={qp.m} nnzA={len(qp.a_val)} nnzQ={len(qp.q_val)} "
        f"in {time.time() - t:.1f}s")
    return bench_single(H, Q, args, qp, args.workload)


def calibration_copy(H):
    """One device-to-device copy of 256 MiB (reads 256 MiB, writes 256 MiB) by
    the library's plain copy kernel: the known-size kernel against which
    FETCH_SIZE / WRITE_SIZE are scaled."""
    lib = H.lib()
    lib.hprqp_hip_calibration_copy.argtypes = [C.c_size_t, C.POINTER(C.c_double)]
    ms = C.c_double(0.0)
    st = lib.hprqp_hip_calibration_copy(256 << 20, C.byref(ms))
    log(f"calibration copy: 268435456 bytes read + written in {ms.value:.3f} ms "
        f"({2 * 268435456 / (ms.value * 1e-3) / 1e9:.0f} GB/s), status {st}")


def calibration_gather(H, nrow=1_000_000, windows=(1_000_000, 500_000, 262_144, 131_072, 4_096)):
    """Random-gather ceiling: an ELL product with 20 coalesced entries per row
    and columns uniform over a window of the gathered vector (the library's
    reference gather kernel), for window sizes from the full 8 MB x down to
    an L1-resident slice."""
    lib = H.lib()
    lib.hprqp_hip_calibration_gather.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    out = []
    for w in windows:
        ms = C.c_double(0.0)
        st = lib.hprqp_hip_calibration_gather(nrow, w, 5, C.byref(ms))
        gathers = 20 * nrow
        byts = 12 * gathers + 8 * nrow
        log(f"gather ceiling: window {w} doubles ({w * 8 / 2**20:.2f} MiB): {ms.value * 1e3:.1f} us for {gathers} "
            f"gathers -> {gathers / (ms.value * 1e-3) / 1e9:.1f} G gathers/s, {byts / (ms.value * 1e-3) / 1e9:.0f} GB/s "
            f"streamed, status {st}")
        out.append(dict(window=w, us=ms.value * 1e3, ggather_s=gathers / (ms.value * 1e-3) / 1e9))
    return out


def profiled_traffic(workload, kernel):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each
    scaled by what the counter reads for the library's plain copy kernel), or
    None when no summary covers it."""
    path = next((p for p in (ROOT / "profiles" / f"{kind}_iteration_kernels_{workload}.md" for kind in ("spmv", "soc"))
                 if p.exists()), None)
    if path is None:
        return None, None
    text = path.read_text()
    start = text.rfind("```json")
    if start < 0:
        return None, None
    data = json.loads(text[start + 7:text.index("```", start + 7)])
    k = data.get("kernels", {}).get(kernel)
    if not k:
        return None, None
    return k["traffic_bytes"], f"profiles/{path.name} (P = {data.get('panels')})"


def bench_single(H, Q, args, qp, workload, soc=False):
    t = time.time()
    if soc:
        model = H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                                  (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u,
                                  soc={"var_offsets": qp.var_offsets})
    else:
        model = build_model(H, qp)
    log(f"host build/preprocess {time.time() - t:.1f}s (nnzQ after symmetrisation {model.c.Q.nnz})")
    if args.profile_only:
        prof = profile_kernels(H, model, args.profile_iters)
        for p in prof:
            log(f"kernel {p['kernel']}: {p['mean_ms'] * 1e3:.1f} us/launch, {p['bytes'] / 1e6:.1f} MB, "
                f"{p['gbs']:.0f} GB/s")
        calibration_copy(H)
        calibration_gather(H)
        return

    for w in range(args.warmup):
        r = H.optimize(model, H.Parameters(verbose=False, max_iter=50, stoptol=args.stoptol))
        log(f"warmup {w}: {r.status} iter={r.iter}")

    results = []
    for k in range(args.steps):
        r = H.optimize(model, H.Parameters(verbose=False, stoptol=args.stoptol))
        log(f"step {k}: {r.status} iter={r.iter} solve={r.time:.3f}s total={r.total_time:.3f}s "
            f"kkt={r.residual:.3e} obj={r.objective:.12g} iter_6={r.iter_6} time_6={r.time_6:.3f}s")
        results.append(r)
    times = [r.time for r in results]
    value = float(np.mean(times))
    iters = int(results[-1].iter)

    prof = profile_kernels(H, model, args.profile_iters)
    for p in prof:
        log(f"kernel {p['kernel']}: {p['mean_ms'] * 1e3:.1f} us/launch, {p['bytes'] / 1e6:.1f} MB, "
            f"{p['gbs']:.0f} GB/s")
    dom = max(prof, key=lambda p: p["mean_ms"] * p["launches"]) if prof else None
    iter_ms = sum(p["mean_ms"] for p in prof)
    iter_bytes = sum(p["bytes"] for p in prof)
    roofline = None
    if dom:
        traffic, source = profiled_traffic(workload, dom["kernel"])
        roofline = {"bound": "hbm", "kernel": dom["kernel"], "achieved": round(dom["gbs"], 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "traffic_source": source,
                    "iteration_gbs": round(iter_bytes / (iter_ms * 1e-3) / 1e9, 1),
                    "iteration_us": round(iter_ms * 1e3, 1)}
        # The SpMV kernels are bound by random gathers, not by streaming: set
        # the dominant kernel against the gather-only reference kernel at the
        # same gather count and window (the gathered vector / column panels).
        s = H.Session(model, H.Parameters(verbose=False))
        # the setup autotuner's choices: column panels and the residency cap
        # (dynamic LDS bytes per workgroup; 0 = uncapped) per matrix
        roofline["panels"] = {m_: int(s.scalar("panels_" + m_)) for m_ in ("Q", "A", "AT")}
        roofline["lds_pad"] = {m_: int(s.scalar("lds_pad_" + m_)) for m_ in ("Q", "A", "AT")}
        mat = {"zxw1_q": "Q", "tempv_q": "Q", "y_a": "A", "w2_at": "AT"}.get(dom["kernel"])
        if mat:
            panels = max(int(s.scalar("panels_" + mat)), 1)
            nnz = {"Q": model.c.Q.nnz, "A": model.c.A.nnz, "AT": model.c.A.nnz}[mat]
            window = (qp.m if mat == "AT" else qp.n) // panels
            g = calibration_gather(H, nrow=max(int(nnz) // 20, 1), windows=(window,))[0]
            roofline["gathers"] = int(nnz)
            roofline["gather_ceiling_us"] = round(g["us"], 1)
            roofline["gather_frac"] = round(g["us"] / (dom["mean_ms"] * 1e3), 3)
        del s

    cpu = None
    if not args.no_cpu:
        s = H.Session(model, H.Parameters(verbose=False))
        cpu = cpu_baseline(qp, iters, s.scalar("lambda_max_A"), s.scalar("lambda_max_Q"),
                           args.cpu_sample if qp.n > 100_000 else 10 ** 9)
        del s
        log("cpu_baseline:", cpu)

    line = {
        "metric": "time-to-1e-8 KKT (plain HPR, 1 GPU)",
        "value": round(value, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(value * 1e3, 2), "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": workload, "n": qp.n, "m": qp.m, "nnzA": int(len(qp.a_val)),
                   "nnzQ": int(model.c.Q.nnz), "stoptol": args.stoptol, "iterations": iters,
                   "status": results[-1].status, "objective": results[-1].objective,
                   "kkt": results[-1].residual, "time_to_1e-6_s": results[-1].time_6,
                   "total_time_s": results[-1].total_time,
                   "us_per_iteration": round(value / max(iters, 1) * 1e6, 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_parallel_solves(Q, qps, stoptol, workers):
    """The oracle (plain-C restatement of the reference CPU path) solving
    `qps` independently on `workers` host threads (the C calls release the
    GIL): returns wall seconds and per-instance statuses."""
    from concurrent.futures import ThreadPoolExecutor

    def one(qp):
        o = Q.OracleSession(qp, stoptol=stoptol)
        return Q.STATUS[o.solve()[0].status]
    t = time.time()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        status = list(ex.map(one, qps))
    return time.time() - t, status


def bench_batched(H, Q, args):
    """Batched workload: args.batch_count QPs with n ~ U{50..2000} in one
    persistent launch (stoptol 1e-6), against sequential plain GPU solves of a
    sample of the same instances."""
    t = time.time()
    qps = Q.batched_instances(count=args.batch_count, seed=0)
    models = [build_model(H, qp) for qp in qps]
    log(f"built {len(models)} instances in {time.time() - t:.1f}s; sum n={sum(q.n for q in qps)}")
    params = H.Parameters(verbose=False, stoptol=1e-6)
    b = H.Batch(models[:8], params)
    b.solve(50)
    del b
    t = time.time()
    b = H.Batch(models, params)
    setup = time.time() - t
    left = b.solve(100)
    secs = b.seconds
    wave_row = b.wave_row
    res = [b.result(k) for k in range(len(models))]
    ok = sum(r.status == "OPTIMAL" for r in res)
    # Algorithmic bytes of one HPR iteration of an instance (general Q,
    # interval rows, between checks): the CSR streams at 12 B per entry (Q
    # twice, A and A^T once each), the row pointers and 23 n + 7 m vector
    # passes; summed over every instance's iterations.
    def iter_bytes(md):
        n, m, nq, na = md.num_variables(), md.num_constraints(), md.c.Q.nnz, md.c.A.nnz
        return 12 * (2 * nq + 2 * na) + 4 * (2 * (n + 1) + (m + 1) + (n + 1)) + 8 * (23 * n + 7 * m)
    inst_iters = sum(int(r.iter) for r in res)
    alg_bytes = sum(int(r.iter) * iter_bytes(md) for r, md in zip(res, models))
    gbs = alg_bytes / secs / 1e9
    prof = ROOT / "profiles" / "batch_solve_batched_4096.md"
    traffic = None
    if prof.exists() and "```json" in prof.read_text():
        text = prof.read_text()
        start = text.rfind("```json")
        traffic = json.loads(text[start + 7:text.index("```", start + 7)])["kernels"]["batch_solve_kernel (CSR)"][
            "traffic_bytes_per_instance_iteration"]
    roofline = {"bound": "hbm", "kernel": "batch_solve_kernel (CSR)", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_unit": "bytes per instance-iteration (profiles/batch_solve_batched_4096.md)",
                "algorithmic_bytes_per_instance_iteration": round(alg_bytes / max(inst_iters, 1)),
                "instance_iterations": inst_iters}
    log(f"roofline: {roofline}")
    # end to end like the sequential solves it is compared with: batch
    # creation (packing, upload, scaling, power iterations) plus the solve
    qps_s = len(models) / (setup + secs)
    log(f"batched: {ok}/{len(models)} optimal, {secs:.3f}s solve + {setup:.3f}s setup -> {qps_s:.1f} QPs/s "
        f"({len(models) / secs:.1f} QPs/s solve only), left={left}")
    sample = list(range(0, len(models), max(1, len(models) // args.seq_sample)))[:args.seq_sample]
    t = time.time()
    seq_iters = 0
    for k in sample:
        r = H.optimize(models[k], params)
        seq_iters += r.iter
    seq = time.time() - t
    seq_qps = len(sample) / seq
    log(f"sequential plain GPU solves: {len(sample)} in {seq:.2f}s -> {seq_qps:.1f} QPs/s")
    cpu = None
    if not args.no_cpu:
        workers = max(1, min(16, os.cpu_count() or 1))
        csample = list(range(0, len(qps), max(1, len(qps) // (4 * workers))))[:4 * workers]
        wall, status = cpu_parallel_solves(Q, [qps[k] for k in csample], 1e-6, workers)
        cpu = dict(value=round(len(csample) / wall, 2), unit="QPs/s", cores=workers, kind="port",
                   sample=f"oracle solves of {len(csample)} of the 4096 instances (evenly spaced), {workers} "
                          f"concurrent host threads, {wall:.2f} s wall, {status.count('OPTIMAL')} optimal")
        log("cpu_baseline:", cpu)
    line = {"metric": "batched QPs/s at 1e-6 (1 GPU)", "value": round(qps_s, 2), "unit": "QPs/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": round((setup + secs) * 1e3, 2), "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator)",
            "config": {"workload": f"batched_{len(models)}", "optimal": ok, "solve_s": secs, "setup_s": setup,
                       "solve_only_qps": round(len(models) / secs, 2),
                       "sequential_plain_gpu_qps": round(seq_qps, 3), "speedup_vs_sequential": round(qps_s / seq_qps, 2),
                       "sequential_sample": len(sample), "wave_row_autotuned": wave_row,
                       "chunk": "100 loop bodies per replay of a captured HIP graph"},
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def bench_layer(H, Q, args):
    """Differentiable layer workload: 1024 dense QPs (n=256, m=320) held as
    FP64 tensors on the GPU (the zero-copy path: no array crosses PCIe);
    forward to 1e-8 and backward (adjoint) times, averaged over --steps
    after --warmup untimed passes. One pass with CPU tensors (host arrays in
    and out) is timed beside it."""
    import torch
    from hprqp_amd import layer as layer_mod
    from hprqp_amd.layer import QPLayer
    count = min(args.batch_count, 1024)
    Qd, c, A, AL, AU, l, u = Q.layer_instances(count)
    dev = torch.device("cuda:0")
    g_host = np.random.default_rng(0).normal(size=(count, Qd.shape[1]))

    def step(where):
        ts = [torch.tensor(a, dtype=torch.float64, device=where, requires_grad=True) for a in (Qd, c, A, AL, AU, l, u)]
        g = torch.tensor(g_host, device=where)
        torch.cuda.synchronize()
        t = time.time()
        x = QPLayer.apply(*ts, 1e-8)
        torch.cuda.synchronize()
        f = time.time() - t
        t = time.time()
        (x * g).sum().backward()
        torch.cuda.synchronize()
        return f, time.time() - t, dict(layer_mod.last_timings)

    for _ in range(args.warmup):
        step(dev)
    runs = [step(dev) for _ in range(max(1, args.steps))]
    fwd = sum(r[0] for r in runs) / len(runs)
    bwd = sum(r[1] for r in runs) / len(runs)
    split = runs[-1][2]
    hf, hb, hsplit = step("cpu")
    log(f"layer (device tensors): forward {fwd:.3f}s backward {bwd:.3f}s; split {split}")
    log(f"layer (host tensors): forward {hf:.3f}s backward {hb:.3f}s; split {hsplit}")
    # Dense column-form products: every instance streams Q once and A twice
    # per iteration (free variables: the noC iteration), 8 bytes per entry;
    # check iterations, setup and vector traffic come on top (not counted).
    n, m = Qd.shape[1], A.shape[1]
    iter_sum = split.get("iterations_sum", 0)
    dsolve = split.get("device_solve", 0.0)
    roofline = None
    if iter_sum and dsolve > 0:
        gbs = 8.0 * (n * n + 2 * m * n) * iter_sum / dsolve / 1e9
        prof = ROOT / "profiles" / "batch_dense_layer_1024.md"
        traffic = None
        if prof.exists():
            text = prof.read_text()
            start = text.rfind("```json")
            if start >= 0:
                k = json.loads(text[start + 7:text.index("```", start + 7)])["kernels"]["batch_solve_kernel (dense)"]
                traffic = k["traffic_bytes_per_instance_iteration"]
        roofline = {"bound": "hbm", "kernel": "batch_solve_kernel (dense)", "achieved": round(gbs, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "traffic_unit": "bytes per instance-iteration (profiles/batch_dense_layer_1024.md)",
                    "bytes_per_instance_iteration": 8 * (n * n + 2 * m * n), "instance_iterations": iter_sum}
    cpu = None
    if not args.no_cpu:
        workers = max(1, min(16, os.cpu_count() or 1))
        csample = list(range(0, count, max(1, count // workers)))[:workers]
        lqps = [Q.qp_from_dense(Qd[k], A[k], c[k], AL[k], AU[k], l[k], u[k]) for k in csample]
        wall, status = cpu_parallel_solves(Q, lqps, 1e-8, workers)
        cpu = dict(value=round(wall * count / len(csample) * 1e3, 1), unit="ms (forward only)",
                   cores=workers, kind="port",
                   sample=f"oracle forward solves (to 1e-8) of {len(csample)} of the {count} instances on {workers} "
                          f"concurrent host threads: {wall:.2f} s wall, {status.count('OPTIMAL')} optimal; value = "
                          f"wall x {count} / {len(csample)} (the all-cores time for the whole batch; no CPU adjoint)")
        log("cpu_baseline:", cpu)
    ms = lambda d: {k: (round(v * 1e3, 2) if not k.startswith("iterations") else v) for k, v in d.items()}
    line = {"metric": "QP layer fwd+bwd ms (1 GPU)", "value": round((fwd + bwd) * 1e3, 2), "unit": "ms",
            "n_gpus": 1, "steps": len(runs), "warmup": args.warmup, "ms_per_step": round((fwd + bwd) * 1e3, 2),
            "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded generator)",
            "config": {"workload": f"layer_{count}", "tensors": "device (zero-copy)", "forward_ms": fwd * 1e3,
                       "backward_ms": bwd * 1e3, "backward_over_forward": bwd / fwd, "split_ms": ms(split),
                       "host_tensors": {"forward_ms": round(hf * 1e3, 2), "backward_ms": round(hb * 1e3, 2),
                                        "split_ms": ms(hsplit)}},
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
