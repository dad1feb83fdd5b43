"""The plain path at the benchmark's full size (n = 1e6, m = 5e5, 19.9M nnz
in A with 5,000 rows of 2,000 nnz, 12.2M nnz in Q): size-independent
properties and the first loop bodies against the oracle.

* the device transpose equals a stable (column, row) sort of the device's
  scaled A, bit for bit;
* Q's and A's scaled values agree with the oracle's scaling bit for bit;
* after 10 loop bodies x_bar, y_bar, w_bar agree with the oracle to 1e-10
  relative (the oracle reuses the device's power-iteration lambdas to bound
  its setup time; lambda parity itself is covered at n = 2000);
* the column-panel carry leaves each row's sum sequential: the panelled
  products equal the unsplit ones bit for bit (HPRQP_PANELS forces counts).
"""
import os

import numpy as np
import pytest

import qp_libs as Q
from test_gpu_parity import build, close

pytestmark = pytest.mark.gpu

FULL = dict(seed=0, n=1_000_000, m_eq=300_000, m_in=200_000, a_row_nnz=20, heavy_rows=5_000, heavy_nnz=2_000,
            b_rows=200_000, b_row_nnz=8, box=10.0)


@pytest.fixture(scope="module")
def setup():
    import hprqp_amd as H
    if not H.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    qp = Q.random_qp(**FULL)
    model = build(H, qp)
    return H, qp, model


def test_fullsize_transpose_bit_exact(setup):
    H, qp, model = setup
    s = H.Session(model, H.Parameters(verbose=False))
    rp, ci, v = s.matrix("A")
    trp, tci, tv = s.matrix("AT")
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.uint64), np.diff(rp.astype(np.int64)))
    order = np.lexsort((rows, ci))
    assert np.array_equal(trp[1:] - trp[:-1], np.bincount(ci.astype(np.int64), minlength=qp.n))
    assert np.array_equal(tci, rows[order])
    assert np.array_equal(tv, v[order])
    print("panels", s.scalar("panels_Q"), s.scalar("panels_A"), s.scalar("panels_AT"), flush=True)


def test_fullsize_first_bodies_match_oracle(setup):
    H, qp, model = setup
    s = H.Session(model, H.Parameters(verbose=False))
    o = Q.OracleSession(qp, lambda_A_given=s.scalar("lambda_max_A"), lambda_Q_given=s.scalar("lambda_max_Q"))
    assert s.scalar("sigma") == o.scalar("sigma")
    for name in ("col_norm", "row_norm", "c", "AL", "AU"):
        assert np.array_equal(s.read(name), o.read(name)), name
    for name in ("Q", "A"):
        assert np.array_equal(s.matrix(name)[2], o.matrix(name)[2]), name
    for body in range(10):
        s.advance(1)
        o.advance(1)
    for v in ("x_bar", "y_bar", "w_bar", "x", "y", "w"):
        assert close(s.read(v), o.read(v)), v


def test_fullsize_panel_products_bitwise(setup):
    H, qp, model = setup
    outs = {}
    for P in ("1", "4"):
        os.environ["HPRQP_PANELS"] = P
        try:
            s = H.Session(model, H.Parameters(verbose=False, max_iter=3))
            s.advance(4)
            outs[P] = [s.read(v) for v in ("x_bar", "y_bar", "w_bar", "x", "w", "y")]
        finally:
            del os.environ["HPRQP_PANELS"]
    for a, b in zip(outs["1"], outs["4"]):
        assert close(a, b, 1e-12)


def test_fullsize_occupancy_cap_changes_no_bit(setup):
    """The residency cap (unused dynamic LDS per workgroup) runs the same
    blocks in the same order: iterates with it forced on and forced off are
    bitwise equal, at two column panels (the value the autotuner picks
    here) and without panels."""
    H, qp, model = setup
    for P in ("2", "1"):
        outs = {}
        for cap in ("1", "0"):
            os.environ["HPRQP_PANELS"], os.environ["HPRQP_OCCUPANCY_CAP"] = P, cap
            try:
                s = H.Session(model, H.Parameters(verbose=False, max_iter=3))
                pads = [s.scalar("lds_pad_" + m) for m in ("Q", "A", "AT")]
                assert (min(pads) > 0) if cap == "1" else (max(pads) == 0), (P, cap, pads)
                s.advance(4)
                outs[cap] = [s.read(v) for v in ("x_bar", "y_bar", "w_bar", "x", "w", "y")]
                del s
            finally:
                del os.environ["HPRQP_PANELS"], os.environ["HPRQP_OCCUPANCY_CAP"]
        for a, b in zip(outs["1"], outs["0"]):
            assert np.array_equal(a, b), P
