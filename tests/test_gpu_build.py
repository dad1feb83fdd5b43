"""build_from_QAbc's preprocessing on the GPU (large models): the
symmetrisation Q <- (Q + Q')/2 with the reference's entry order and 1e-15
drop rule (src/shared/build.c:157-252) and the transposes of A
(src/shared/sparse.c:81-125), against the oracle's sequential restatement
and against this library's host-thread path (HPRQP_HOST_BUILD=1), bit for
bit; on unsorted and repeated entries, entries at the drop tolerance, pairs
that cancel exactly, empty and free rows of A, and the n = 1e6 workload."""
import os
import time

import numpy as np
import pytest

import qp_libs as Q
from test_abi import big_messy_qp, model_arrays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible")
    return hprqp_amd


def build(H, qp):
    return H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                             (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u)


def host_build(H, qp):
    os.environ["HPRQP_HOST_BUILD"] = "1"
    try:
        return build(H, qp)
    finally:
        del os.environ["HPRQP_HOST_BUILD"]


def with_cancelling_pairs(qp, seed, pairs=50_000):
    """Appends to random rows r an entry (r, c, v) and to row c an entry
    (c, r, -v): their contributions to Q_sym[r][c] cancel exactly, so the
    entry must disappear (unless other entries share the position). Free
    rows (AL = -inf, AU = +inf) are added to A."""
    rng = np.random.default_rng(seed)
    n = qp.n
    rows = np.repeat(np.arange(n), np.diff(qp.q_ptr.astype(np.int64)))
    r = rng.integers(0, n, size=pairs)
    c = rng.integers(0, n, size=pairs)
    v = rng.standard_normal(pairs)
    all_r = np.concatenate([rows, r, c])
    all_c = np.concatenate([qp.q_ind.astype(np.int64), c, r])
    all_v = np.concatenate([qp.q_val, v, -v])
    order = np.argsort(all_r, kind="stable")  # rows grouped, entry order kept within a row
    qp.q_ind = all_c[order].astype(np.uint64)
    qp.q_val = all_v[order]
    qp.q_ptr = np.concatenate([[0], np.cumsum(np.bincount(all_r, minlength=n))]).astype(np.uint64)
    AL, AU = qp.AL.copy(), qp.AU.copy()
    free = rng.random(qp.m) < 0.01
    AL[free], AU[free] = -np.inf, np.inf
    qp.AL, qp.AU = AL, AU
    return qp


@pytest.mark.parametrize("seed", [7, 8])
def test_device_build_matches_oracle_and_host_path(H, seed, capfd):
    qp = with_cancelling_pairs(big_messy_qp(seed=seed), seed)
    print(f"device build seed={seed}: nnzQ={len(qp.q_val)} nnzA={len(qp.a_val)}", flush=True)
    os.environ["HPRQP_BUILD_TRACE"] = "1"
    try:
        dev = model_arrays(build(H, qp))
    finally:
        del os.environ["HPRQP_BUILD_TRACE"]
    err = capfd.readouterr().err
    assert "Q symmetrised on the device" in err and "A transposed on the device" in err, err
    host = model_arrays(host_build(H, qp))
    o = Q.OracleSession(qp, use_scaling=0, max_iter=0)
    assert o.m < qp.m  # empty and free rows were deleted
    for d3, h3, name in zip(dev, host, ("Q", "A", "AT")):
        for d, h, w_ in zip(d3, h3, o.matrix(name)):
            assert np.array_equal(d, w_), name
            assert np.array_equal(d, h), name


def test_device_build_of_the_large_workload(H, capfd):
    qp = Q.random_qp(**Q.PLAIN_N1M)
    os.environ["HPRQP_BUILD_TRACE"] = "1"
    try:
        t = time.time()
        dev = model_arrays(build(H, qp))
        t_dev = time.time() - t
    finally:
        del os.environ["HPRQP_BUILD_TRACE"]
    assert "Q symmetrised on the device" in capfd.readouterr().err
    t = time.time()
    host = model_arrays(host_build(H, qp))
    t_host = time.time() - t
    print(f"plain_n1M build: device {t_dev:.3f} s, host threads {t_host:.3f} s", flush=True)
    for d3, h3, name in zip(dev, host, ("Q", "A", "AT")):
        for d, h in zip(d3, h3):
            assert np.array_equal(d, h), name
