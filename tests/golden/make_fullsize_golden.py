"""Writes the full-size fixtures of the plain and batched workloads from the
reference CPU path built from its own sources (oracle/_ref/libhprqp_ref.so,
made by oracle/Makefile.ref) and the C oracle (oracle/liboracle.so):

* plain_n1M.json (python make_fullsize_golden.py plain):
  the n = 1e6, m = 5e5 instance (A: 300k equality + 200k two-sided rows at
  20 nnz, 5,000 rows at 2,000 nnz; Q = B'DB + 1e-2 I, B 200k x n at 8 nnz),
  stoptol 1e-8, defaults otherwise, no polishing:
  - sha256 of the generated arrays (the GPU test regenerates and checks it);
  - the reference's full solve: status, iterations, objective, KKT residual,
    gap, iter_6 (first iteration below 1e-6) and its solve-loop time;
  - the oracle's lambda_max(AA') and lambda_max(Q) with their power-iteration
    step counts (src/cpu/eigs.c:19-185) and sigma0;
  - the oracle's iterates after loop bodies 10 and 100: 4,096 sampled entries
    of x_bar, y_bar, w_bar and each vector's 2-norm;
  - the oracle's own full solve (status, iterations, objective), which must
    equal the reference's bit for bit.
* soc_n220k.json (python make_fullsize_golden.py soc): the same records for
  the cone workload (20,000 ten-dimensional cones after a 20,000-variable
  box prefix, 100,000 equality rows at 40 nnz, diagonal Q) from the oracle
  alone -- the reference has no CPU cone path (src/shared/solver.c:250-253)
  -- plus the independent certificate of the oracle's solution
  (tests/gpu/test_gpu_soc.cu:152-196).
* batched_4096.npz (python make_fullsize_golden.py batched):
  the 4,096 batched instances (n ~ U{50..2000}, m = round(n/2)) solved one by
  one by the reference at stoptol 1e-6: status, iterations, objective, KKT.

Run once in a container that has /root/reference (about 10 minutes on 8
cores). The GPU tests only read the files written here.
"""
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import qp_libs as Q  # noqa: E402
from fullsize import qp_sha256  # noqa: E402

PLAIN = dict(seed=0, n=1_000_000, m_eq=300_000, m_in=200_000, a_row_nnz=20, heavy_rows=5_000, heavy_nnz=2_000,
             b_rows=200_000, b_row_nnz=8, box=10.0)
SAMPLE = 1024
BODIES = (10, 100)


def sample_indices(length, seed):
    # every fourth of 4,096 sorted random positions
    return np.sort(np.random.default_rng(seed).choice(length, size=min(4 * SAMPLE, length), replace=False))[::4]


def single(name, qp, stoptol=1e-8, reference=True):
    """Fixture of one large instance; see the module docstring."""
    out = {"workload": name, "sha256": qp_sha256(qp), "stoptol": stoptol}
    t = time.time()
    if reference:
        r = Q.binding_solve(Q.ref_lib(), qp, use_gpu=0, stoptol=stoptol)
        out["reference"] = {k: r[k] for k in ("status", "iter", "objective", "residual", "gap", "iter_6", "time")}
        print("reference", out["reference"], f"{time.time() - t:.1f}s", flush=True)
    t = time.time()
    o = Q.OracleSession(qp, stoptol=stoptol)
    out["oracle_setup"] = {k: o.scalar(k) for k in ("lambda_max_A", "lambda_max_Q", "power_steps_A",
                                                     "power_steps_Q", "sigma", "Q_is_diag", "noC")}
    print("oracle setup", out["oracle_setup"], f"{time.time() - t:.1f}s", flush=True)
    done = 0
    iterates = {}
    for b in BODIES:
        o.advance(b - done)
        done = b
        rec = {}
        for v in ("x_bar", "y_bar", "w_bar"):
            a = o.read(v)
            idx = sample_indices(len(a), {"x_bar": 1, "y_bar": 2, "w_bar": 3}[v])
            rec[v] = {"norm": float(np.linalg.norm(a)), "index": idx.tolist(), "value": a[idx].tolist()}
        rec["sigma"] = o.scalar("sigma")
        iterates[str(b)] = rec
    out["oracle_iterates"] = iterates
    res, x, y, z, w = o.solve()
    out["oracle_full"] = {"status": Q.STATUS[res.status], "iter": int(res.iter), "objective": res.primal_obj,
                          "residual": res.residuals, "iter_6": int(res.iter_6)}
    if qp.num_var_cones or qp.num_row_cones:
        out["oracle_certificate"] = {k: float(v) for k, v in Q.soc_certificate(qp, x, y, z).items()}
    print("oracle full", out["oracle_full"], f"{time.time() - t:.1f}s", flush=True)
    if reference:
        assert out["oracle_full"]["status"] == out["reference"]["status"]
        assert out["oracle_full"]["iter"] == out["reference"]["iter"]
        assert out["oracle_full"]["objective"] == out["reference"]["objective"]
    return out


def plain():
    t = time.time()
    qp = Q.random_qp(**PLAIN)
    print(f"generated in {time.time() - t:.1f}s", flush=True)
    out = single("plain_n1M", qp)
    out["generator"] = PLAIN
    (HERE / "plain_n1M.json").write_text(json.dumps(out, separators=(",", ":")) + "\n")


def soc():
    qp = Q.soc_qp(**Q.SOC_N220K)
    out = single("soc_n220k", qp, reference=False)
    out["generator"] = Q.SOC_N220K
    (HERE / "soc_n220k.json").write_text(json.dumps(out, separators=(",", ":")) + "\n")


def _solve_one(k):
    qp = _INSTANCES[k]
    r = Q.binding_solve(Q.ref_lib(), qp, use_gpu=0, stoptol=1e-6)
    return Q.STATUS_CODE[r["status"]], r["iter"], r["objective"], r["residual"]


_INSTANCES = None


def batched():
    global _INSTANCES
    t = time.time()
    _INSTANCES = Q.batched_instances(count=4096, seed=0)
    sha = hashlib.sha256()
    for qp in _INSTANCES:
        sha.update(qp_sha256(qp).encode())
    workers = min(8, os.cpu_count() or 1)
    with ProcessPoolExecutor(max_workers=workers) as ex:
        rows = list(ex.map(_solve_one, range(len(_INSTANCES)), chunksize=16))
    st = np.array([r[0] for r in rows], np.int32)
    np.savez_compressed(HERE / "batched_4096.npz", status=st, iter=np.array([r[1] for r in rows], np.int64),
                        objective=np.array([r[2] for r in rows]), kkt=np.array([r[3] for r in rows]),
                        sha256=np.array(sha.hexdigest()), stoptol=np.array(1e-6))
    print(f"batched: {int((st == Q.STATUS_CODE['OPTIMAL']).sum())} optimal of {len(st)} in {time.time() - t:.1f}s")


if __name__ == "__main__":
    which = sys.argv[1:] or ["batched", "soc", "plain"]
    if "batched" in which:
        batched()
    if "soc" in which:
        soc()
    if "plain" in which:
        plain()
