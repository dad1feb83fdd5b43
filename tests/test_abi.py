"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares; host-side model building (no device) matches the
oracle's preprocessing bit for bit."""
import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import qp_libs as Q

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "hpr-qp-c_amd" / "libhprqp_hip.so"


def declared_functions():
    names = set()
    for h in (ROOT / "include").rglob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//.*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}()]*)\)\s*;", text):
            name = m.group(1)
            if name not in {"if", "while", "for", "return", "sizeof"}:
                names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "libhprqp_hip.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = declared_functions()
    assert len(declared) > 60
    missing = [n for n in declared if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_no_device_here():
    import hprqp_amd  # noqa: F401  (loads it after PyTorch's HIP runtime)
    lib = C.CDLL(str(LIB))
    lib.hprqp_status_name.restype = C.c_char_p
    assert lib.hprqp_status_name(6) == b"OPTIMAL"
    assert lib.hprqp_status_name(99) == b"UNKNOWN"


def test_package_has_no_cpu_fallback():
    import hprqp_amd as H
    if H.gpu_backend_available():
        pytest.skip("a device is present")
    m = H.build_from_qabc([[2.0, 0.0], [0.0, 2.0]], [-4.0, -2.0], [], [], [], [-math.inf] * 2, [math.inf] * 2)
    r = H.optimize(m, H.Parameters(verbose=False))
    assert r.status == "NOT_IMPLEMENTED"
    r = H.optimize(m, H.Parameters(verbose=False, use_gpu=False))
    assert r.status == "NOT_IMPLEMENTED"


def model_arrays(model):
    def csr(M):
        rp = np.ctypeslib.as_array(M.row_ptr, shape=(M.nrow + 1,)).copy()
        ci = np.ctypeslib.as_array(M.col_ind, shape=(max(M.nnz, 1),))[:M.nnz].copy()
        v = np.ctypeslib.as_array(M.values, shape=(max(M.nnz, 1),))[:M.nnz].copy()
        return rp, ci, v
    return csr(model.c.Q), csr(model.c.A), csr(model.c.AT)


@pytest.mark.parametrize("kw", [dict(seed=0), dict(seed=7, n=300, m_eq=90, m_in=60, a_row_nnz=9, heavy_rows=2,
                                                   heavy_nnz=280, b_rows=80, b_row_nnz=6)])
def test_host_build_matches_oracle_preprocess(kw):
    import hprqp_amd as H
    qp = Q.random_qp(**kw)
    # perturb Q so symmetrisation has real work: Q' = Q + upper-triangular noise
    model = H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                              (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u)
    (qrp, qci, qv), (arp, aci, av), (trp, tci, tv) = model_arrays(model)
    o = Q.OracleSession(qp, use_scaling=0, max_iter=0)
    for got, name in (((qrp, qci, qv), "Q"), ((arp, aci, av), "A"), ((trp, tci, tv), "AT")):
        want = o.matrix(name)
        for g, w_ in zip(got, want):
            assert np.array_equal(g, w_), name


def test_host_build_rejects_invalid_input():
    import hprqp_amd as H
    with pytest.raises(ValueError):
        H.build_from_qabc([[1.0]], [0.0], [[1.0]], [1.0], [0.0], [-1.0], [1.0])  # AL > AU
    with pytest.raises(ValueError):
        H.build_from_qabc([[math.nan]], [0.0], [[1.0]], [0.0], [1.0], [-1.0], [1.0])
    with pytest.raises(ValueError):
        H.build_from_qabc([[1.0]], [0.0], [[1.0]], [0.0], [1.0], [2.0], [1.0])  # l > u


def test_host_build_symmetrises_and_drops_rows():
    import hprqp_amd as H
    inf = math.inf
    Qm = (2, 2, np.array([0, 2, 3], np.uint64), np.array([0, 1, 1], np.uint64), np.array([2.0, 6.0, 4.0]))
    Am = (3, 2, np.array([0, 2, 2, 2], np.uint64), np.array([0, 1], np.uint64), np.array([1.0, -1.0]))
    m = H.build_from_qabc(Qm, [0.0, 0.0], Am, [0.0, -inf, 1.0], [3.0, inf, 2.0], [-2.0, -2.0], [2.0, 2.0])
    (qrp, qci, qv), (arp, aci, av), _ = model_arrays(m)
    assert m.num_constraints() == 1 and list(qv) == [2.0, 3.0, 3.0, 4.0]


def messy_q_qp(seed=3, n=60, m=20):
    """A QP whose Q is unsymmetric, with unsorted columns, repeated
    (row, col) entries (diagonal ones included) and entries at and below the
    1e-15 drop tolerance, so the symmetrisation order and drop rules matter."""
    rng = np.random.default_rng(seed)
    ptr, ind, val = [0], [], []
    for r in range(n):
        cols = list(rng.choice(n, size=rng.integers(0, 8), replace=True))
        if rng.random() < 0.6:
            cols += [r] * int(rng.integers(1, 3))
        rng.shuffle(cols)
        for cidx in cols:
            ind.append(int(cidx))
            v = rng.standard_normal() * 10.0 ** rng.integers(-3, 3)
            if rng.random() < 0.1:
                v = rng.choice([1e-15, -2e-15, 1.5e-15, 3e-16])
            val.append(float(v))
        ptr.append(len(ind))
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.2)
    A[0] = 0.0  # an empty row, dropped by preprocessing
    base = Q.qp_from_dense(np.eye(n), A, rng.standard_normal(n), -np.ones(m), np.ones(m), -np.ones(n), np.ones(n))
    base.q_ptr = np.array(ptr, np.uint64)
    base.q_ind = np.array(ind, np.uint64)
    base.q_val = np.array(val, np.float64)
    return base


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_host_build_matches_oracle_on_messy_q(seed):
    import hprqp_amd as H
    qp = messy_q_qp(seed)
    model = H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                              (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u)
    got = model_arrays(model)
    o = Q.OracleSession(qp, use_scaling=0, max_iter=0)
    for g3, name in zip(got, ("Q", "A", "AT")):
        for g, w_ in zip(g3, o.matrix(name)):
            assert np.array_equal(g, w_), name


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("seed", [3, 4, 5])
def test_oracle_symmetrisation_equals_reference_on_messy_q(seed):
    # The reference's own preprocessing of the same messy Q, seen through its
    # solve: identical iterates after 40 iterations need identical matrices.
    qp = messy_q_qp(seed)
    r = Q.binding_solve(Q.ref_lib(), qp, use_gpu=0, max_iter=40)
    o = Q.OracleSession(qp, max_iter=40)
    res, x, y, z, w = o.solve()
    assert res.primal_obj == r["objective"]
    for v, arr in zip("xyzw", (x, y, z, w)):
        assert np.array_equal(arr, r[v]), v


def big_messy_qp(seed=7, n=200_000, m=60_000, q_row=14, a_row=40):
    """messy_q_qp's pattern at a size where the host build runs its
    transposes and the symmetrisation on worker threads (several million
    entries): unsorted and repeated Q entries, entries around the 1e-15 drop
    tolerance, a few empty A rows."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 2 * q_row, size=n)
    q_ptr = np.zeros(n + 1, np.uint64)
    q_ptr[1:] = np.cumsum(lens)
    nq = int(q_ptr[-1])
    q_ind = rng.integers(0, n, size=nq).astype(np.uint64)
    rows = np.repeat(np.arange(n), lens)
    diag = rng.random(nq) < 0.08
    q_ind[diag] = rows[diag]  # repeated diagonal entries
    q_val = rng.standard_normal(nq) * 10.0 ** rng.integers(-3, 3, size=nq)
    tiny = rng.random(nq) < 0.05
    q_val[tiny] = rng.choice([1e-15, -2e-15, 1.5e-15, 3e-16], size=int(tiny.sum()))
    alen = rng.integers(0, 2 * a_row, size=m)
    alen[rng.random(m) < 0.01] = 0
    a_ptr = np.zeros(m + 1, np.uint64)
    a_ptr[1:] = np.cumsum(alen)
    na = int(a_ptr[-1])
    a_ind = np.empty(na, np.uint64)
    for r in range(m):  # sorted, distinct columns per row
        b, e = int(a_ptr[r]), int(a_ptr[r + 1])
        a_ind[b:e] = np.sort(rng.choice(n, size=e - b, replace=False))
    a_val = rng.standard_normal(na)
    return Q.QP(n, m, q_ptr, q_ind, q_val, a_ptr, a_ind, a_val, rng.standard_normal(n), -np.ones(m), np.ones(m),
                -np.ones(n), np.ones(n))


def test_threaded_host_build_matches_oracle():
    """The worker-thread transposes and symmetrisation give the oracle's
    (the reference's sequential) model bit for bit."""
    import hprqp_amd as H
    qp = big_messy_qp()
    assert len(qp.q_val) > (1 << 21) and len(qp.a_val) > (1 << 21)
    model = H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                              (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val), qp.AL, qp.AU, qp.l, qp.u)
    got = model_arrays(model)
    o = Q.OracleSession(qp, use_scaling=0, max_iter=0)
    for g3, name in zip(got, ("Q", "A", "AT")):
        for g, w_ in zip(g3, o.matrix(name)):
            assert np.array_equal(g, w_), name


def _ref_build(qp):
    """build_from_QAbc of the reference library (oracle/_ref): (status, QP)."""
    L = Q.ref_lib()
    keep = []

    def mat(nrow, ncol, rp, ci, v):
        rp, ci, v = (np.ascontiguousarray(a if len(a) else np.zeros(1, a.dtype)) for a in
                     (np.asarray(rp, np.uint64), np.asarray(ci, np.uint64), np.asarray(v, np.float64)))
        keep.extend([rp, ci, v])
        return Q._CSR(nrow, ncol, int(rp[-1]) if nrow else 0, rp.ctypes.data_as(Q.c_size_p),
                      ci.ctypes.data_as(Q.c_size_p), v.ctypes.data_as(Q.c_double_p), 0)

    def dv(a):
        a = np.ascontiguousarray(a if len(a) else np.zeros(1), np.float64)
        keep.append(a)
        return a.ctypes.data_as(Q.c_double_p)
    L.build_from_QAbc.argtypes = [C.POINTER(Q._CSR), C.POINTER(Q._CSR)] + [Q.c_double_p] * 5 + [C.c_double,
                                                                                                C.POINTER(Q._CModel)]
    m = Q._CModel()
    Qm = mat(qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val)
    Am = mat(qp.m, qp.n, qp.a_ptr if qp.m else np.zeros(1, np.uint64), qp.a_ind, qp.a_val)
    st = L.build_from_QAbc(C.byref(Qm), C.byref(Am), dv(qp.c), dv(qp.AL), dv(qp.AU), dv(qp.l), dv(qp.u),
                           C.c_double(qp.obj_constant), C.byref(m))
    try:
        return st, (Q.qp_of_cmodel(m) if st == 0 else None)
    finally:
        L.model_free(C.byref(m))


def _random_messy(rng, n, m):
    """A small model with some of: NaN or inf entries, crossed bounds, free
    and empty rows, repeated and unsorted entries, sub-tolerance values."""
    def csr(nrow, ncol, per_row):
        ptr, ind, val = [0], [], []
        for _ in range(nrow):
            k = int(rng.integers(0, per_row + 1))
            cols = list(rng.integers(0, ncol, size=k))
            rng.shuffle(cols)
            ind += [int(c) for c in cols]
            val += [float(v) for v in rng.standard_normal(k) * 10.0 ** rng.integers(-2, 2, size=k)]
            ptr.append(len(ind))
        return np.array(ptr, np.uint64), np.array(ind, np.uint64), np.array(val)
    qptr, qind, qval = csr(n, n, 4)
    aptr, aind, aval = csr(m, n, 4)
    c = rng.standard_normal(n)
    lo = rng.standard_normal(m) - 1.0
    AL, AU = lo, lo + rng.uniform(0, 2, size=m)
    l, u = -np.ones(n) * 3, np.ones(n) * 3
    for _ in range(int(rng.integers(0, 3))):  # damage
        kind = int(rng.integers(8))
        if kind == 0 and len(qval):
            qval[int(rng.integers(len(qval)))] = rng.choice([np.nan, np.inf, 1e-16, -3e-16])
        elif kind == 1 and len(aval):
            aval[int(rng.integers(len(aval)))] = rng.choice([np.nan, -np.inf, 0.0])
        elif kind == 2 and m:
            i = int(rng.integers(m))
            AL[i], AU[i] = AU[i] + 1.0, AL[i]  # crossed
        elif kind == 3 and m:
            i = int(rng.integers(m))
            AL[i], AU[i] = -np.inf, np.inf  # free row: deleted
        elif kind == 4:
            j = int(rng.integers(n))
            l[j], u[j] = 2.0, 1.0  # crossed
        elif kind == 5:
            c[int(rng.integers(n))] = rng.choice([np.nan, np.inf])
        elif kind == 6 and m:
            AL[int(rng.integers(m))] = np.nan
        elif kind == 7:
            j = int(rng.integers(n))
            l[j], u[j] = -np.inf, np.inf
    return Q.QP(n, m, qptr, qind, qval, aptr, aind, aval, c, AL, AU, l, u, float(rng.standard_normal()))


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
def test_random_builds_agree_with_reference():
    """build_from_QAbc of this library and of the reference on 300 small
    random models, a share of them invalid: the same status, and bitwise
    equal models when both accept."""
    import hprqp_amd as H
    rng = np.random.default_rng(11)
    ok = 0
    for it in range(300):
        n, m = int(rng.integers(1, 9)), int(rng.integers(0, 7))
        qp = _random_messy(rng, n, m)
        st_ref, ref = _ref_build(qp)
        try:
            model = H.build_from_qabc((qp.n, qp.n, qp.q_ptr, qp.q_ind, qp.q_val), qp.c,
                                      (qp.m, qp.n, qp.a_ptr, qp.a_ind, qp.a_val) if qp.m else np.zeros((0, qp.n)),
                                      qp.AL, qp.AU, qp.l, qp.u, qp.obj_constant)
            mine = Q.qp_of_cmodel(model.c)
        except ValueError:
            mine = None
        assert (mine is None) == (st_ref != 0), (it, st_ref)
        if mine is not None:
            Q.assert_same_model(mine, ref)
            ok += 1
    assert 60 < ok < 300


CLIENT = ROOT / "tests" / "c_client"
REF_INCLUDE = Path("/root/reference/include")


def _compile(src, out, include, libs=()):
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{include}", str(src), "-o", str(out),
                    *libs], check=True, capture_output=True, text=True)


def test_abi_struct_layouts_match_reference_headers(tmp_path):
    """sizeof / alignof / offsetof of every field of the ABI structs of
    include/shared/types.h and the enum values, compiled against this
    repository's headers and against the reference's
    (include/shared/types.h:10-214): identical, so a client compiled against
    either runs with this library."""
    if not REF_INCLUDE.is_dir():
        pytest.skip("the reference's headers are not present here")
    outs = []
    for k, inc in enumerate((ROOT / "include", REF_INCLUDE)):
        exe = tmp_path / f"layout{k}"
        _compile(CLIENT / "abi_layout.c", exe, inc)
        outs.append(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    assert outs[0].count("\n") > 140
    assert outs[0] == outs[1]


def test_c_client_compiles_against_include_and_links():
    """A plain C11 client (tests/c_client/client.c) compiles with -Wall
    -Werror against include/ alone and links to libhprqp_hip.so; run here
    without a device, each solve reports NOT_IMPLEMENTED (no CPU fallback)
    and the CPU path is refused."""
    subprocess.run(["make", "-s", "-C", str(CLIENT)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(CLIENT / "_build" / "client"), str(ROOT / "tests" / "golden" / "mps" / "simple_qp.mps")],
                         check=True, capture_output=True, text=True).stdout
    lines = {ln.split()[0]: dict(kv.split("=", 1) for kv in ln.split()[1:]) for ln in out.splitlines()}
    assert lines["done"]["rc"] == "0"
    assert lines["binding_mps_read"] == {"rc": "OK", "n": "2", "m": "2"}
    assert lines["model_api_cpu"]["rc"] == "NOT_IMPLEMENTED"
    import hprqp_amd as H
    if not H.gpu_backend_available():
        for case in ("binding_unconstrained", "binding_mps_max_iter", "model_api_equality"):
            assert lines[case]["rc"] == "NOT_IMPLEMENTED", case
