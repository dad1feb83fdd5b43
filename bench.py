"""Benchmark of the HPR-QP iteration on one MI355X (FP64).

Headline (the metric value): time to KKT < 1e-8 of the plain solver at scale,
plain_n1M -- a seeded random sparse convex QP with n = 1,000,000 variables,
m = 500,000 rows (300k equality, 200k two-sided), A at 20 nnz/row with 1% of
rows (5,000) at 2,000 nnz, Q = B'DB + 1e-2 I with B 200,000 x n at 8 nnz/row,
box [-10, 10], Ruiz scaling on, check_iter 100, print_frequency -1, no
polishing, no Gram. A "step" is one full solve from a fresh device session
(upload, device transpose, scaling, power iterations) to stoptol; the value
is the solve-loop time (the reference's results.time), lower is better.
--warmup W runs W short solves (max_iter 50) first. Inputs are host arrays
handed over through the C ABI, as the reference's callers do; the solve-loop
time starts with the data resident on the device.

The default run (--workload all) also measures, as sub-objects of the same
JSON line ("workloads"), the other workloads the metric names:
  plain_n2k     time to 1e-8 of the n = 2,000 instance, against the
                reference CPU path's full solve on one host core
  soc_n220k     time to 1e-8 of the cone workload (20,000 cones of dimension
                10 after a 20,000-variable box prefix, 100,000 equality rows)
  batched_4096  QPs/s at 1e-6 of 4,096 QPs (n ~ U{50..2000}) in one batch,
                against sequential plain GPU solves and host-core solves
  layer_1024    forward + backward ms of the differentiable layer on 1,024
                dense QPs (n = 256, m = 320) held as ROCm tensors
  gram_grid63   time to 1e-8 of the grid-control model (g = 63) with the
                equality-row (Gram) metric that AUTO takes, against the same
                model with the metric off
Each carries its own roofline and cpu_baseline; --workload NAME runs one.

  roofline      the dominant kernel's algorithmic bytes per launch / its mean
                launch time (HIP events on the library's own stream), against
                MI355X's 8 TB/s HBM peak; traffic is the counter-measured HBM
                bytes per launch from the committed rocprofv3 summary
  cpu_baseline  the reference CPU path (oracle/_ref, built from the
                reference's sources) where a whole solve fits the time budget,
                else the oracle (the plain-C restatement of that path, one
                core) on a bounded sample of the same workload, scaled to the
                GPU solve's iteration count; host model and core count stated
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "hpr-qp-c_amd"))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def _workloads():
    import qp_libs as Q
    return {"plain_n1M": Q.PLAIN_N1M, "plain_n2k": Q.PLAIN_N2K}


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


def build_model(H, qp):
    soc = None
    if qp.num_var_cones or qp.num_row_cones:
        soc = {}
        if qp.num_var_cones:
            soc["var_offsets"] = qp.var_offsets
        if qp.num_row_cones:
            soc["row_offsets"] = qp.row_offsets
            soc["rhs"] = qp.soc_rhs
    return H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                             (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u,
                             qp.obj_constant, soc=soc)


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


def _profile_json(name):
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    text = path.read_text()
    start = text.rfind("```json")
    if start < 0:
        return None
    return json.loads(text[start + 7:text.index("```", start + 7)])


def profiled_traffic(workload, kernel):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each
    scaled by what the counter reads for the library's plain copy kernel), or
    None when no summary covers it."""
    for kind in ("spmv", "soc"):
        name = f"{kind}_iteration_kernels_{workload}.md"
        data = _profile_json(name)
        if data and data.get("kernels", {}).get(kernel):
            return data["kernels"][kernel]["traffic_bytes"], f"profiles/{name} (P = {data.get('panels')})"
    return None, None


# ----------------------------------------------------------------- CPU baselines

def cpu_oracle_sample(qp, gpu_iters, lam_a, lam_q, bodies, stoptol=1e-8):
    """The oracle (one core) on loop bodies 1..`bodies` of the workload after
    its setup (the power iterations skipped: the lambdas are given), scaled
    to the GPU solve's iteration count."""
    import qp_libs as Q
    t0 = time.time()
    o = Q.OracleSession(qp, stoptol=stoptol, lambda_A_given=lam_a, lambda_Q_given=lam_q)
    setup = time.time() - t0
    o.advance(1)                      # body 0: iteration-0 residuals + first step
    t1 = time.time()
    o.advance(bodies)
    dt = time.time() - t1
    per_iter = dt / bodies
    return dict(value=round(per_iter * gpu_iters, 3), unit="s (est. solve loop to 1e-8)", cores=1, kind="port",
                sample=f"oracle (C restatement of the reference CPU path), loop bodies 1..{bodies} of this workload "
                       f"after setup ({setup:.1f} s, power iterations skipped with the lambdas given): "
                       f"{per_iter * 1e3:.2f} ms/iteration x {gpu_iters} iterations",
                host=host_info())


def cpu_full_solve(qp, stoptol=1e-8):
    """The whole solve on one host core: the reference CPU path built from its
    sources (oracle/_ref) through its binding API (solve_qp --use-gpu 0),
    else the oracle."""
    import qp_libs as Q
    if Q.ref_available():
        t0 = time.time()
        r = Q.binding_solve(Q.ref_lib(), qp, use_gpu=0, stoptol=stoptol)
        dt = time.time() - t0
        return dict(value=r["time"], unit="s (solve loop to 1e-8)", cores=1, kind="reference",
                    sample=f"reference CPU path (oracle/_ref/libhprqp_ref.so), whole solve ({r['iter']} iterations, "
                           f"{r['status']}); setup + solve {dt:.3f} s", host=host_info())
    t0 = time.time()
    r = Q.OracleSession(qp, stoptol=stoptol).solve()[0]
    dt = time.time() - t0
    return dict(value=r.time, unit="s (solve loop to 1e-8)", cores=1, kind="port",
                sample=f"oracle, whole solve ({r.iter} iterations, {Q.STATUS[r.status]}); setup + solve {dt:.3f} s",
                host=host_info())


def cpu_parallel_solves(Q, qps, stoptol, workers):
    """The oracle solving `qps` independently on `workers` host threads (the C
    calls release the GIL): returns wall seconds and per-instance statuses."""
    from concurrent.futures import ThreadPoolExecutor

    def one(qp):
        o = Q.OracleSession(qp, stoptol=stoptol)
        return Q.STATUS[o.solve()[0].status]
    t = time.time()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        status = list(ex.map(one, qps))
    return time.time() - t, status


# ----------------------------------------------------------------- single QP

def bench_single(H, Q, args, qp, workload, steps, warmup, cpu_bodies, gather=True):
    t = time.time()
    model = build_model(H, qp)
    log(f"{workload}: model build (build_from_QAbc) {time.time() - t:.2f}s (nnzQ after symmetrisation {model.c.Q.nnz})")
    if args.profile_only:
        prof = profile_kernels(H, model, args.profile_iters)
        for p in prof:
            log(f"kernel {p['kernel']}: {p['mean_ms'] * 1e3:.1f} us/launch, {p['bytes'] / 1e6:.1f} MB, "
                f"{p['gbs']:.0f} GB/s")
        calibration_copy(H)
        calibration_gather(H)
        # each iteration kernel against the gather ceiling at its own gather
        # count and window
        s = H.Session(model, H.Parameters(verbose=False))
        mats = {"zxw1_q": "Q", "tempv_q": "Q", "y_a": "A", "w2_at": "AT"}
        for p in prof:
            mat = mats.get(p["kernel"])
            if not mat:
                continue
            panels = max(int(s.scalar("panels_" + mat)), 1)
            nnz = {"Q": model.c.Q.nnz, "A": model.c.A.nnz, "AT": model.c.A.nnz}[mat]
            window = (qp.m if mat == "AT" else qp.n) // panels
            g = calibration_gather(H, nrow=max(int(nnz) // 20, 1), windows=(window,))[0]
            log(f"kernel {p['kernel']}: {nnz} gathers over a {window}-double window (P = {panels}, "
                f"lds_pad {int(s.scalar('lds_pad_' + mat))}): ceiling {g['us']:.1f} us, kernel "
                f"{p['mean_ms'] * 1e3:.1f} us -> gather_frac {g['us'] / (p['mean_ms'] * 1e3):.3f}")
        return None

    for w in range(warmup):
        r = H.optimize(model, H.Parameters(verbose=False, max_iter=50, stoptol=args.stoptol))
        log(f"{workload} warmup {w}: {r.status} iter={r.iter}")
    results = []
    for k in range(steps):
        r = H.optimize(model, H.Parameters(verbose=False, stoptol=args.stoptol))
        log(f"{workload} step {k}: {r.status} iter={r.iter} solve={r.time:.4f}s total={r.total_time:.3f}s "
            f"kkt={r.residual:.3e} obj={r.objective:.12g} iter_6={r.iter_6} time_6={r.time_6:.4f}s")
        results.append(r)
    value = float(np.mean([r.time for r in results]))
    iters = int(results[-1].iter)

    prof = profile_kernels(H, model, args.profile_iters)
    for p in prof:
        log(f"{workload} kernel {p['kernel']}: {p['mean_ms'] * 1e3:.1f} us/launch, {p['bytes'] / 1e6:.1f} MB, "
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
                    "iteration_us": round(iter_ms * 1e3, 1),
                    "kernels_us": {p["kernel"]: round(p["mean_ms"] * 1e3, 2) for p in prof}}
        s = H.Session(model, H.Parameters(verbose=False))
        # the setup autotuner's choices: column panels and the residency cap
        # (dynamic LDS bytes per workgroup; 0 = uncapped) per matrix
        roofline["panels"] = {m_: int(s.scalar("panels_" + m_)) for m_ in ("Q", "A", "AT")}
        roofline["lds_pad"] = {m_: int(s.scalar("lds_pad_" + m_)) for m_ in ("Q", "A", "AT")}
        lam_a, lam_q = s.scalar("lambda_max_A"), s.scalar("lambda_max_Q")
        if gather:
            # The SpMV kernels are bound by random gathers, not by streaming:
            # each iteration kernel against the gather-only reference kernel
            # at the same gather count and window (the gathered vector or its
            # column panel).
            mats = {"zxw1_q": "Q", "tempv_q": "Q", "y_a": "A", "w2_at": "AT"}
            fracs = {}
            for p in prof:
                mat = mats.get(p["kernel"])
                if not mat:
                    continue
                panels = max(int(s.scalar("panels_" + mat)), 1)
                nnz = {"Q": model.c.Q.nnz, "A": model.c.A.nnz, "AT": model.c.A.nnz}[mat]
                window = (qp.m if mat == "AT" else qp.n) // panels
                g = calibration_gather(H, nrow=max(int(nnz) // 20, 1), windows=(window,))[0]
                fracs[p["kernel"]] = {"gather_ceiling_us": round(g["us"], 1), "kernel_us": round(p["mean_ms"] * 1e3, 1),
                                      "gather_frac": round(g["us"] / (p["mean_ms"] * 1e3), 3)}
            roofline["gather"] = fracs
            if dom["kernel"] in fracs:
                roofline["gather_frac"] = fracs[dom["kernel"]]["gather_frac"]
        del s
    else:
        s = H.Session(model, H.Parameters(verbose=False))
        lam_a, lam_q = s.scalar("lambda_max_A"), s.scalar("lambda_max_Q")
        del s

    cpu = None
    if not args.no_cpu:
        cpu = cpu_full_solve(qp, args.stoptol) if cpu_bodies is None else \
            cpu_oracle_sample(qp, iters, lam_a, lam_q, cpu_bodies, args.stoptol)
        cpu["speedup"] = round(cpu["value"] / value, 2) if value > 0 else None
        log(f"{workload} cpu_baseline:", cpu)
    return {"value": round(value, 5), "unit": "s", "ms_per_step": round(value * 1e3, 3), "steps": steps,
            "config": {"workload": workload, "n": qp.n, "m": qp.m, "nnzA": int(len(qp.a_val)),
                       "nnzQ": int(model.c.Q.nnz), "cones": qp.num_var_cones, "stoptol": args.stoptol,
                       "iterations": iters, "status": results[-1].status, "objective": results[-1].objective,
                       "kkt": results[-1].residual, "time_to_1e-6_s": results[-1].time_6,
                       "total_time_s": results[-1].total_time,
                       "us_per_iteration": round(value / max(iters, 1) * 1e6, 2)},
            "roofline": roofline, "cpu_baseline": cpu}


# ----------------------------------------------------------------- batched

def bench_batched(H, Q, args):
    """args.batch_count QPs with n ~ U{50..2000} in one persistent launch
    (stoptol 1e-6), end to end (batch creation included), against sequential
    plain GPU solves of a sample of the same instances and host-core solves."""
    t = time.time()
    qps = Q.batched_instances(count=args.batch_count, seed=0)
    models = [build_model(H, qp) for qp in qps]
    log(f"batched: built {len(models)} instances in {time.time() - t:.1f}s; sum n={sum(q.n for q in qps)}")
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
    info = b.status()
    ok = sum(s_ == "OPTIMAL" for s_ in info["status"])
    del b

    # Algorithmic bytes of one HPR iteration of an instance (general Q,
    # interval rows, between checks): the CSR streams at 12 B per entry (Q
    # twice, A and A^T once each), the row pointers and 23 n + 7 m vector
    # passes; summed over every instance's iterations.
    def iter_bytes(md):
        n, m, nq, na = md.num_variables(), md.num_constraints(), md.c.Q.nnz, md.c.A.nnz
        return 12 * (2 * nq + 2 * na) + 4 * (2 * (n + 1) + (m + 1) + (n + 1)) + 8 * (23 * n + 7 * m)
    its = [int(v) for v in info["iter"]]
    inst_iters = sum(its)
    alg_bytes = sum(k * iter_bytes(md) for k, md in zip(its, models))
    gbs = alg_bytes / secs / 1e9
    data = _profile_json("batch_solve_batched_4096.md")
    traffic = data["kernels"]["batch_solve_kernel (CSR)"]["traffic_bytes_per_instance_iteration"] if data else None
    roofline = {"bound": "hbm", "kernel": "batch_solve_kernel (CSR)", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_unit": "bytes per instance-iteration (profiles/batch_solve_batched_4096.md)",
                "algorithmic_bytes_per_instance_iteration": round(alg_bytes / max(inst_iters, 1)),
                "instance_iterations": inst_iters}
    qps_s = len(models) / (setup + secs)
    log(f"batched: {ok}/{len(models)} optimal, {secs:.3f}s solve + {setup:.3f}s setup -> {qps_s:.1f} QPs/s "
        f"({len(models) / secs:.1f} QPs/s solve only), left={left}")
    sample = list(range(0, len(models), max(1, len(models) // args.seq_sample)))[:args.seq_sample]
    t = time.time()
    for k in sample:
        H.optimize(models[k], params)
    seq = time.time() - t
    seq_qps = len(sample) / seq
    log(f"batched: sequential plain GPU solves: {len(sample)} in {seq:.2f}s -> {seq_qps:.1f} QPs/s")
    cpu = None
    if not args.no_cpu:
        workers = max(1, min(16, os.cpu_count() or 1))
        csample = list(range(0, len(qps), max(1, len(qps) // (4 * workers))))[:4 * workers]
        wall, status = cpu_parallel_solves(Q, [qps[k] for k in csample], 1e-6, workers)
        cpu = dict(value=round(len(csample) / wall, 2), unit="QPs/s", cores=workers, kind="port",
                   sample=f"oracle solves of {len(csample)} of the {len(qps)} instances (evenly spaced), {workers} "
                          f"concurrent host threads, {wall:.2f} s wall, {status.count('OPTIMAL')} optimal",
                   host=host_info(), speedup=round(qps_s / (len(csample) / wall), 2))
        log("batched cpu_baseline:", cpu)
    return {"value": round(qps_s, 2), "unit": "QPs/s", "ms_per_step": round((setup + secs) * 1e3, 2),
            "config": {"workload": f"batched_{len(models)}", "optimal": ok, "solve_s": round(secs, 4),
                       "setup_s": round(setup, 4), "solve_only_qps": round(len(models) / secs, 2),
                       "sequential_plain_gpu_qps": round(seq_qps, 3), "speedup_vs_sequential": round(qps_s / seq_qps, 2),
                       "sequential_sample": len(sample), "wave_row_autotuned": wave_row, "stoptol": 1e-6,
                       "chunk": "100 loop bodies per replay of a captured HIP graph"},
            "roofline": roofline, "cpu_baseline": cpu}


# ----------------------------------------------------------------- layer

def bench_gram(H, Q, args, steps, warmup, g=63):
    """The equality-row (Gram) metric on the grid-control model at g = 63
    (3,969 metric rows, n = 4,473, m = 4,725, 252 row cones), the size of the
    reference's live AUTO case: time to 1e-8 with the metric (AUTO takes it)
    against the same model with the metric off; the roofline of the metric's
    solve (two triangular matrix-vector products over W = L^-1 and W', each
    e(e+1)/2 entries, streamed from HBM)."""
    qp = Q.gram_fixture_spatial(g)
    model = build_model(H, qp)
    if args.profile_only:
        for p in profile_kernels(H, model, args.profile_iters):
            log(f"kernel {p['kernel']}: {p['mean_ms'] * 1e3:.1f} us/launch, {p['bytes'] / 1e6:.2f} MB, "
                f"{p['gbs']:.0f} GB/s")
        return None
    for w in range(warmup):
        H.optimize(model, H.Parameters(verbose=False, max_iter=50, stoptol=args.stoptol))
    results = []
    for k in range(steps):
        r = H.optimize(model, H.Parameters(verbose=False, stoptol=args.stoptol))
        log(f"gram_grid{g} step {k}: {r.status} gram_active={r.gram_active} iter={r.iter} solve={r.time:.4f}s "
            f"total={r.total_time:.3f}s kkt={r.residual:.3e} obj={r.objective:.12g}")
        results.append(r)
    off = H.optimize(model, H.Parameters(verbose=False, stoptol=args.stoptol, gram_mode=0, max_iter=200000))
    log(f"gram_grid{g} metric off: {off.status} iter={off.iter} solve={off.time:.4f}s obj={off.objective:.12g}")
    value = float(np.mean([r.time for r in results]))
    prof = profile_kernels(H, model, args.profile_iters)
    for p in prof:
        log(f"gram_grid{g} kernel {p['kernel']}: {p['mean_ms'] * 1e3:.1f} us/launch, {p['bytes'] / 1e6:.2f} MB, "
            f"{p['gbs']:.0f} GB/s")
    roofline = None
    gram = [p for p in prof if p["kernel"] == "gram_solve"]
    if gram:
        p = gram[0]
        data = _profile_json(f"gram_solve_gram_grid{g}.md")
        roofline = {"bound": "hbm", "kernel": "gram_solve (W and W' triangular products + y finish)",
                    "achieved": round(p["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(p["gbs"] / HBM_PEAK_GBS, 4),
                    "traffic": data.get("traffic_bytes_per_step") if data else None,
                    "traffic_source": f"profiles/gram_solve_gram_grid{g}.md" if data else None,
                    "kernel_us": round(p["mean_ms"] * 1e3, 2),
                    "iteration_us": round(sum(q["mean_ms"] for q in prof) * 1e3, 1),
                    "kernels_us": {q["kernel"]: round(q["mean_ms"] * 1e3, 2) for q in prof}}
    cpu = None
    if not args.no_cpu:
        # the oracle (one core) with the same metric: its setup (a dense
        # Cholesky of G) and then loop bodies 1..100, scaled to the GPU's
        # iteration count
        t0 = time.time()
        o = Q.OracleSession(qp, stoptol=args.stoptol)
        setup = time.time() - t0
        o.advance(1)
        t1 = time.time()
        o.advance(100)
        per = (time.time() - t1) / 100
        iters = int(results[-1].iter)
        cpu = dict(value=round(per * iters, 3), unit="s (est. solve loop to 1e-8)", cores=1, kind="port",
                   sample=f"oracle with the Gram metric (dense Cholesky of G, {setup:.1f} s setup), loop bodies "
                          f"1..100: {per * 1e3:.2f} ms/iteration x {iters} iterations", host=host_info())
        cpu["speedup"] = round(cpu["value"] / value, 2) if value > 0 else None
        log(f"gram_grid{g} cpu_baseline:", cpu)
    r = results[-1]
    return {"value": round(value, 5), "unit": "s", "ms_per_step": round(value * 1e3, 3), "steps": steps,
            "config": {"workload": f"gram_grid{g}", "n": qp.n, "m": qp.m, "metric_rows": g * g,
                       "row_cones": qp.num_row_cones, "gram_active": bool(r.gram_active), "stoptol": args.stoptol,
                       "iterations": int(r.iter), "status": r.status, "objective": r.objective, "kkt": r.residual,
                       "metric_off": {"status": off.status, "iterations": int(off.iter), "solve_s": round(off.time, 4),
                                      "objective": off.objective},
                       "speedup_vs_metric_off": round(off.time / value, 2) if value > 0 else None},
            "roofline": roofline, "cpu_baseline": cpu}


def bench_layer(H, Q, args, steps, warmup):
    """1024 dense QPs (n=256, m=320) held as FP64 tensors on the GPU (the
    zero-copy path: no array crosses PCIe); forward to 1e-8 and backward
    (adjoint) times, averaged over `steps` after `warmup` untimed passes. One
    pass with CPU tensors (host arrays in and out) is timed beside it."""
    import torch
    from hprqp_amd import layer as layer_mod
    from hprqp_amd.layer import QPLayer
    count = min(args.batch_count, 1024)
    Qd, c, A, AL, AU, l, u = Q.layer_instances(count)
    dev = torch.device("cuda:0")
    g_host = np.random.default_rng(12345).normal(size=(count, Qd.shape[1]))

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

    for _ in range(warmup):
        step(dev)
    runs = [step(dev) for _ in range(max(1, steps))]
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
        data = _profile_json("batch_dense_layer_1024.md")
        traffic = data["kernels"]["batch_solve_kernel (dense)"]["traffic_bytes_per_instance_iteration"] if data else None
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
                          f"wall x {count} / {len(csample)} (the all-cores time for the whole batch; no CPU adjoint)",
                   host=host_info(), speedup_forward=round(wall * count / len(csample) / fwd, 2))
        log("layer cpu_baseline:", cpu)
    ms = lambda d: {k: (round(v * 1e3, 3) if not k.startswith("iterations") and "residual" not in k else v)
                    for k, v in d.items()}
    return {"value": round((fwd + bwd) * 1e3, 2), "unit": "ms", "ms_per_step": round((fwd + bwd) * 1e3, 2),
            "steps": len(runs),
            "config": {"workload": f"layer_{count}", "tensors": "device (zero-copy)", "forward_ms": round(fwd * 1e3, 2),
                       "backward_ms": round(bwd * 1e3, 2), "backward_over_forward": round(bwd / fwd, 3),
                       "stoptol": 1e-8, "split_ms": ms(split),
                       "host_tensors": {"forward_ms": round(hf * 1e3, 2), "backward_ms": round(hb * 1e3, 2),
                                        "split_ms": ms(hsplit)}},
            "roofline": roofline, "cpu_baseline": cpu}


# ----------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all",
                    choices=["all", "plain_n1M", "plain_n2k", "plain_n1M_banded", "plain_n1M_no_heavy",
                             "plain_n1M_only_heavy", "soc_n220k", "batched_4096", "layer_1024", "gram_grid63"])
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--stoptol", type=float, default=1e-8)
    ap.add_argument("--profile-iters", type=int, default=20)
    ap.add_argument("--cpu-bodies", type=int, default=300,
                    help="loop bodies of the plain_n1M CPU sample (about 70 ms each on one core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="set up the workload, run --profile-iters timed iterations and a 256 MiB device copy "
                         "(the reference kernel for counter calibration); no solve, no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the solver runs on one GPU, and "
              "the multi-GPU sharding of this path is design-only.", file=sys.stderr)
        sys.exit(2)

    import hprqp_amd as H
    import qp_libs as Q
    if not H.gpu_backend_available():
        raise SystemExit("no gfx950 device visible")
    log("device:", H.device_name(), "host:", host_info())
    common = {"n_gpus": 1, "warmup": args.warmup, "vs_baseline": None, "dtype": "f64",
              "data": "synthetic (seeded generator)"}

    def emit(metric, res, higher):
        line = {"metric": metric, "value": res["value"], "unit": res["unit"], "n_gpus": 1,
                "steps": res.get("steps", 1), "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
                "higher_is_better": higher, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
                "data": common["data"], "config": res["config"], "roofline": res["roofline"],
                "cpu_baseline": res["cpu_baseline"]}
        return line

    w = args.workload
    if w == "batched_4096":
        print(json.dumps(emit("batched QPs/s at 1e-6 (1 GPU)", bench_batched(H, Q, args), True)), flush=True)
        return
    if w == "layer_1024":
        print(json.dumps(emit("QP layer fwd+bwd ms (1 GPU)", bench_layer(H, Q, args, args.steps, args.warmup),
                              False)), flush=True)
        return
    if w == "gram_grid63":
        res = bench_gram(H, Q, args, args.steps, args.warmup)
        if res:
            print(json.dumps(emit("time-to-1e-8 KKT (Gram metric, 1 GPU)", res, False)), flush=True)
        return
    if w == "soc_n220k":
        qp = Q.soc_qp(**Q.SOC_N220K)
        res = bench_single(H, Q, args, qp, "soc_n220k", args.steps, args.warmup, 100, gather=False)
        if res:
            print(json.dumps(emit("time-to-1e-8 KKT (cone HPR, 1 GPU)", res, False)), flush=True)
        return
    if w.startswith("plain_"):
        variant = w[len("plain_n1M"):] if w.startswith("plain_n1M_") else ""
        if variant and not args.profile_only:
            raise SystemExit(f"{w} is a profiling aid: use it with --profile-only")
        kw = dict(_workloads()["plain_n2k" if w == "plain_n2k" else "plain_n1M"])
        if variant == "_no_heavy":
            # only the 20-nnz rows of A (stream blocks): 500,000 rows, 10M nnz
            kw.update(heavy_rows=0)
        elif variant == "_only_heavy":
            # only 2,000-nnz rows (wave blocks): 5,000 rows, 10M nnz
            kw.update(m_eq=3_000, m_in=2_000, a_row_nnz=2_000, heavy_rows=0)
        t = time.time()
        qp = Q.random_qp(**kw)
        if variant == "_banded":
            qp = banded_variant(qp)
        log(f"generated {w}: n={qp.n} m={qp.m} nnzA={len(qp.a_val)} in {time.time() - t:.1f}s")
        small = w == "plain_n2k"
        res = bench_single(H, Q, args, qp, w, max(args.steps, 5) if small else args.steps, args.warmup,
                           None if small else args.cpu_bodies, gather=not small)
        if res:
            print(json.dumps(emit("time-to-1e-8 KKT (plain HPR, 1 GPU)", res, False)), flush=True)
        return

    # all: the headline workload, then the others as sub-objects of one line
    t = time.time()
    qp = Q.random_qp(**Q.PLAIN_N1M)
    log(f"generated plain_n1M: n={qp.n} m={qp.m} nnzA={len(qp.a_val)} in {time.time() - t:.1f}s")
    head = bench_single(H, Q, args, qp, "plain_n1M", args.steps, args.warmup, args.cpu_bodies)
    del qp
    line = emit("time-to-1e-8 KKT (plain HPR, 1 GPU)", head, False)
    subs = {}
    qp = Q.random_qp(**Q.PLAIN_N2K)
    r = bench_single(H, Q, args, qp, "plain_n2k", 5, 1, None, gather=False)
    subs["plain_n2k"] = {"metric": "time-to-1e-8 KKT (plain HPR, n = 2,000, 1 GPU)", "higher_is_better": False, **r}
    qp = Q.soc_qp(**Q.SOC_N220K)
    r = bench_single(H, Q, args, qp, "soc_n220k", 3, 1, 100, gather=False)
    subs["soc_n220k"] = {"metric": "time-to-1e-8 KKT (cone HPR, 1 GPU)", "higher_is_better": False, **r}
    del qp
    r = bench_batched(H, Q, args)
    subs["batched_4096"] = {"metric": "batched QPs/s at 1e-6 (1 GPU)", "higher_is_better": True, **r}
    r = bench_layer(H, Q, args, 3, 1)
    subs["layer_1024"] = {"metric": "QP layer fwd+bwd ms (1 GPU)", "higher_is_better": False, **r}
    r = bench_gram(H, Q, args, 3, 1)
    subs["gram_grid63"] = {"metric": "time-to-1e-8 KKT (Gram metric, 1 GPU)", "higher_is_better": False, **r}
    line["workloads"] = subs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
