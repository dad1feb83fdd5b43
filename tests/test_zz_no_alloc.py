"""hprqp_hip_csr_spmv_noalloc (the SpMV entry point that allocates nothing:
its row-block plan lives in the library's fixed device array) against
hprqp_hip_csr_spmv, through the C ABI on device buffers: bit-identical y on
two calls in a row, the second with a smaller matrix, so a stale plan from
the first call would show; sizes beyond the stated limit are rejected."""
import ctypes as C

import numpy as np
import pytest

import qp_libs as Q

pytestmark = pytest.mark.gpu

LIMIT = 1 << 19


@pytest.fixture(scope="module")
def lib():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible: the HIP path must run on the GPU")
    L = hprqp_amd.lib()
    for f in (L.hprqp_hip_csr_spmv, L.hprqp_hip_csr_spmv_noalloc):
        f.argtypes = [C.c_int] + [C.c_void_p] * 6
    return L


def both(L, lengths, ncol, seed):
    rp, ci, val = Q.csr_with_row_lengths(lengths, ncol, seed=seed)
    x = np.random.default_rng(seed + 1).standard_normal(ncol)
    bufs = [Q.DeviceArray(rp.astype(np.int32)), Q.DeviceArray(ci.astype(np.int32)), Q.DeviceArray(val),
            Q.DeviceArray(x)]
    outs = [Q.DeviceArray(shape=len(lengths), dtype=np.float64) for _ in range(2)]
    try:
        print(f"spmv noalloc vs plain: nrow={len(lengths)} nnz={rp[-1]} seed={seed}", flush=True)
        st = L.hprqp_hip_csr_spmv(len(lengths), *[b.ptr for b in bufs], outs[0].ptr, None)
        assert st == 0, Q.STATUS.get(st, st)
        st = L.hprqp_hip_csr_spmv_noalloc(len(lengths), *[b.ptr for b in bufs], outs[1].ptr, None)
        assert st == 0, Q.STATUS.get(st, st)
        return outs[0].get(), outs[1].get()
    finally:
        for b in bufs + outs:
            b.close()


def test_noalloc_matches_allocating_entry_point_twice(lib):
    rng = np.random.default_rng(7)
    # first call: long enough for several plan chunks and row-pointer windows,
    # with stream, wave and single-row blocks mixed
    big = rng.choice([0, 1, 5, 20, 256, 257, 600, 3000, 5000], size=40_000,
                     p=[0.05, 0.1, 0.3, 0.45, 0.02, 0.02, 0.03, 0.02, 0.01])
    y_ref, y_new = both(lib, big, 50_000, seed=11)
    np.testing.assert_array_equal(y_new, y_ref)
    # second call: a smaller matrix (fewer blocks than the plan left behind)
    small = rng.choice([0, 3, 17, 300], size=700)
    y_ref, y_new = both(lib, small, 900, seed=12)
    np.testing.assert_array_equal(y_new, y_ref)


def test_noalloc_limits(lib):
    assert lib.hprqp_hip_csr_spmv_noalloc(0, None, None, None, None, None, None) == 0
    # rejected before any device access: the pointers are never read
    assert lib.hprqp_hip_csr_spmv_noalloc(LIMIT + 1, 8, 8, 8, 8, 8, None) == Q.STATUS_CODE["INVALID_ARGUMENT"]
