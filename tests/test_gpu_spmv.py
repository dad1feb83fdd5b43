"""Kernel-level parity of the row-blocked SpMV (hprqp_hip_csr_spmv, device
pointers through the C ABI) with the oracle's sequential CSR product
(or_csr_matvec, the reference's src/cpu/sparse_ops.c row loop).

Rows of at most 256 entries are summed left to right in LDS without FMA, so
they must match bit for bit. Longer rows (one wavefront per row up to 4096
entries, the whole workgroup beyond) use FMA and a tree: they are held to
4 * nnz_row * eps relative to sum |a_ij x_j|, the bound for any summation
order.
"""
import ctypes as C

import numpy as np
import pytest

import qp_libs as Q

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible: the HIP path must run on the GPU")
    L = hprqp_amd.lib()
    L.hprqp_hip_csr_spmv.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
    return L


def oracle_matvec(rp, ci, val, x):
    y = np.zeros(len(rp) - 1)
    args = [Q._sz(rp), Q._sz(ci), Q._db(val), Q._db(x)]
    Q.oracle_lib().or_csr_matvec(len(rp) - 1, *[p for _, p in args], y.ctypes.data_as(Q.c_double_p))
    return y


def device_matvec(L, rp, ci, val, x):
    bufs = [Q.DeviceArray(rp.astype(np.int32)), Q.DeviceArray(ci.astype(np.int32)), Q.DeviceArray(val),
            Q.DeviceArray(x), Q.DeviceArray(shape=len(rp) - 1, dtype=np.float64)]
    try:
        print(f"csr_spmv nrow={len(rp) - 1} nnz={rp[-1]}", flush=True)
        st = L.hprqp_hip_csr_spmv(len(rp) - 1, *[b.ptr for b in bufs], None)
        assert st == 0, Q.STATUS.get(st, st)
        return bufs[-1].get()
    finally:
        for b in bufs:
            b.close()


LENGTH_CASES = {
    # stream blocks: short rows, empty rows, a block cut by the 2048-entry cap
    "short": [3, 0, 17, 256, 1, 0, 0, 255] * 40 + [0] * 7,
    "cap": [250] * 30,
    # wave blocks: runs of 1..5 long rows between short ones, at the 256/257 edge
    "wave": [257, 5, 300, 301, 302, 303, 304, 2, 4096, 4095, 1000, 0, 257, 258],
    # vector blocks: single rows beyond 4096, next to wave rows
    "vector": [4097, 9000, 3000, 20000, 1, 4097],
    "mixed": list(np.random.default_rng(5).choice([0, 1, 7, 40, 256, 257, 700, 2000, 5000], size=600)),
}


@pytest.mark.parametrize("case", sorted(LENGTH_CASES))
def test_spmv_matches_sequential_rows(lib, case):
    lengths = np.asarray(LENGTH_CASES[case])
    ncol = 30000
    rp, ci, val = Q.csr_with_row_lengths(lengths, ncol, seed=1)
    x = np.random.default_rng(2).standard_normal(ncol)
    y_dev = device_matvec(lib, rp, ci, val, x)
    y_ref = oracle_matvec(rp, ci, val, x)
    short = lengths <= 256
    np.testing.assert_array_equal(y_dev[short], y_ref[short])
    if (~short).any():
        absum = np.array([np.abs(val[rp[r]:rp[r + 1]] * x[ci[rp[r]:rp[r + 1]]]).sum() for r in range(len(lengths))])
        bound = 4 * lengths * np.finfo(float).eps * absum
        err = np.abs(y_dev - y_ref)
        assert (err[~short] <= bound[~short]).all(), float((err / np.maximum(bound, 1e-300))[~short].max())


def test_spmv_empty_matrix(lib):
    assert lib.hprqp_hip_csr_spmv(0, None, None, None, None, None, None) == 0
