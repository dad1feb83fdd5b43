"""Writes tests/golden/reference_traces.npz from the reference CPU path built
from its own sources (oracle/_ref/libhprqp_ref.so, made by oracle/Makefile.ref).

For each reference unit-test fixture (tests/test_support.c:79-270) and two
small seeded random QPs it records the reported results of a solve stopped at
max_iter = k (k in KS) and of the full solve: status, iteration count,
objective, KKT residual, gap and the unscaled x, y, z, w. Those are the
reference's own outputs; the oracle and the HIP path are checked against them.

Run: python tests/golden/make_golden.py   (needs /root/reference)
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import qp_libs as Q  # noqa: E402

KS = [1, 2, 3, 5, 12, 50, 99, 100, 101, 150, 300]
RANDOM = {"rand60": dict(seed=1, n=60, m_eq=20, m_in=10, a_row_nnz=5, b_rows=15, b_row_nnz=3),
          "rand200": dict(seed=2, n=200, m_eq=60, m_in=40, a_row_nnz=8, heavy_rows=2, heavy_nnz=100,
                          b_rows=50, b_row_nnz=4)}
SKIP_FULL = {"TALL_DIAG"}  # infeasible: runs to max_iter = INT_MAX


def main():
    L = Q.ref_lib()
    cases = dict(Q.reference_fixtures())
    for name, kw in RANDOM.items():
        cases[name] = Q.random_qp(**kw)
    out = {}
    for name, qp in cases.items():
        for k in KS + [None]:
            if k is None and name in SKIP_FULL:
                continue
            kw = {} if k is None else {"max_iter": k}
            r = Q.binding_solve(L, qp, use_gpu=0, **kw)
            tag = f"{name}/{'full' if k is None else k}"
            out[tag + "/scalars"] = np.array([Q.STATUS_CODE[r["status"]], r["iter"], r["objective"],
                                              r["residual"], r["gap"]], dtype=np.float64)
            for v in "xyzw":
                out[tag + "/" + v] = r[v]
    np.savez_compressed(HERE / "reference_traces.npz", **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
