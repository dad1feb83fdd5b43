"""ctypes access to the three libraries the tests compare.

* oracle  -- oracle/liboracle.so, the CPU restatement (test infrastructure).
* ref     -- oracle/_ref/libhprqp_ref.so, the reference CPU path compiled from
             its own sources (present only where it was built).
* gen     -- hpr-qp-c_amd/libhprqp_gen.so, the seeded workload generators.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "hpr-qp-c_amd"

c_size_p = C.POINTER(C.c_size_t)
c_double_p = C.POINTER(C.c_double)

STATUS = {0: "OK", 1: "INVALID_ARGUMENT", 2: "ALLOCATION_FAILED", 3: "NOT_IMPLEMENTED",
          4: "MAX_ITER", 5: "TIME_LIMIT", 6: "OPTIMAL", 7: "CONTINUE"}
STATUS_CODE = {v: k for k, v in STATUS.items()}


def _sz(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(c_size_p)


def _db(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(c_double_p)


@dataclass
class QP:
    """A QP in the reference's build_from_QAbc form (CSR, size_t indices)."""
    n: int
    m: int
    q_ptr: np.ndarray
    q_ind: np.ndarray
    q_val: np.ndarray
    a_ptr: np.ndarray
    a_ind: np.ndarray
    a_val: np.ndarray
    c: np.ndarray
    AL: np.ndarray
    AU: np.ndarray
    l: np.ndarray
    u: np.ndarray
    obj_constant: float = 0.0
    var_offsets: np.ndarray | None = None
    row_offsets: np.ndarray | None = None
    soc_rhs: np.ndarray | None = None
    x0: np.ndarray | None = None

    @property
    def num_var_cones(self):
        return 0 if self.var_offsets is None else len(self.var_offsets) - 1

    @property
    def num_row_cones(self):
        return 0 if self.row_offsets is None else len(self.row_offsets) - 1

    def dense_Q(self):
        Q = np.zeros((self.n, self.n))
        for r in range(self.n):
            for k in range(self.q_ptr[r], self.q_ptr[r + 1]):
                Q[r, self.q_ind[k]] += self.q_val[k]
        return Q

    def dense_A(self):
        A = np.zeros((self.m, self.n))
        for r in range(self.m):
            for k in range(self.a_ptr[r], self.a_ptr[r + 1]):
                A[r, self.a_ind[k]] += self.a_val[k]
        return A


def qp_from_dense(Q, A, c, AL, AU, l, u, obj_constant=0.0):
    Q = np.atleast_2d(np.asarray(Q, dtype=float))
    n = Q.shape[1]
    A = np.asarray(A, dtype=float).reshape(-1, n)

    def csr(M):
        ptr, ind, val = [0], [], []
        for row in M:
            for j, v in enumerate(row):
                if v != 0.0:
                    ind.append(j)
                    val.append(v)
            ptr.append(len(ind))
        return (np.array(ptr, dtype=np.uint64), np.array(ind, dtype=np.uint64), np.array(val, dtype=float))

    qp, qi, qv = csr(Q)
    ap, ai, av = csr(A)
    return QP(n, A.shape[0], qp, qi, qv, ap, ai, av, np.asarray(c, float), np.asarray(AL, float),
              np.asarray(AU, float), np.asarray(l, float), np.asarray(u, float), obj_constant)


def reference_fixtures():
    """The reference's 12 unit-test models (tests/test_support.c:79-270)."""
    inf = math.inf
    F = {}
    F["STANDARD"] = qp_from_dense([[2, 0], [0, 4]], [[1, 1]], [2, 1], [2], [2], [0, 0], [10, 10])
    F["STANDARD_NOC"] = qp_from_dense([[2, 0], [0, 4]], [[1, 1]], [2, 1], [2], [2], [-inf, -inf], [inf, inf])
    F["SCALE"] = qp_from_dense([[4, 0], [0, 9]], [[16, 25]], [2, 4], [2], [6], [0, -1], [10, 3])
    F["RESIDUALS"] = qp_from_dense([[2, 0], [0, 4]], [[1, 2]], [1, -1], [0], [100], [0, 0], [10, 10])
    F["DENSE_Q"] = qp_from_dense([[2, 1], [1, 2]], [[1, 2]], [0, 0], [-1], [1], [-10, -10], [10, 10])
    F["CONTROL_NONDIAG"] = qp_from_dense([[2, .5], [.5, 3]], [[1, -1]], [1, -.5], [0], [1.5], [-10, -10], [10, 10])
    F["EMPTY_A"] = qp_from_dense([[2, 0], [0, 3]], np.zeros((0, 2)), [1, -.5], [], [], [-inf, -inf], [inf, inf])
    F["TALL_DIAG"] = qp_from_dense([[3, 0], [0, 5]], [[1, 1], [1, -1], [2, .5]], [1.5, -.5], [1, -.5, .25],
                                   [1, -.5, .25], [0, -1], [2, 3])
    F["DISPATCH_1D"] = qp_from_dense([[1]], np.zeros((0, 1)), [0], [], [], [-1], [1])
    F["GPU_SPMV_1D"] = qp_from_dense([[2]], [[3]], [0], [-1], [1], [-10], [10])
    F["GPU_ITERATION_STANDARD"] = qp_from_dense([[2, .5], [.5, 3]], [[1, -2]], [.5, -.25], [-.5], [.75],
                                                [-1, -1], [1, 1])
    F["ZERO_1D"] = qp_from_dense(np.zeros((1, 1)), np.zeros((0, 1)), [0], [], [], [-1], [1])
    return F


# --------------------------------------------------------------------------- generator

class _GenProblem(C.Structure):
    _fields_ = [("n", C.c_size_t), ("m", C.c_size_t), ("m_eq", C.c_size_t),
                ("q_nnz", C.c_size_t), ("a_nnz", C.c_size_t),
                ("q_ptr", c_size_p), ("q_ind", c_size_p), ("q_val", c_double_p),
                ("a_ptr", c_size_p), ("a_ind", c_size_p), ("a_val", c_double_p),
                ("c", c_double_p), ("AL", c_double_p), ("AU", c_double_p), ("l", c_double_p),
                ("u", c_double_p), ("x0", c_double_p),
                ("num_var_cones", C.c_size_t), ("var_offsets", c_size_p),
                ("num_row_cones", C.c_size_t), ("row_offsets", c_size_p), ("soc_rhs", c_double_p)]


_gen = None


def gen_lib():
    global _gen
    if _gen is None:
        path = PKG / "libhprqp_gen.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        _gen = C.CDLL(str(path))
        _gen.qpgen_random_qp.argtypes = [C.c_uint64] + [C.c_size_t] * 8 + [C.c_double, C.POINTER(_GenProblem)]
        _gen.qpgen_soc_qp.argtypes = [C.c_uint64] + [C.c_size_t] * 5 + [C.c_double, C.POINTER(_GenProblem)]
        _gen.qpgen_free.argtypes = [C.POINTER(_GenProblem)]
    return _gen


def _take(gp: _GenProblem) -> QP:
    def arr(p, n, dt):
        if n == 0 or not p:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)
    n, m = gp.n, gp.m
    qp = QP(n, m, arr(gp.q_ptr, n + 1, np.uint64), arr(gp.q_ind, gp.q_nnz, np.uint64), arr(gp.q_val, gp.q_nnz, float),
            arr(gp.a_ptr, m + 1, np.uint64), arr(gp.a_ind, gp.a_nnz, np.uint64), arr(gp.a_val, gp.a_nnz, float),
            arr(gp.c, n, float), arr(gp.AL, m, float), arr(gp.AU, m, float), arr(gp.l, n, float),
            arr(gp.u, n, float), 0.0, x0=arr(gp.x0, n, float))
    if gp.num_var_cones:
        qp.var_offsets = arr(gp.var_offsets, gp.num_var_cones + 1, np.uint64)
    if gp.num_row_cones:
        qp.row_offsets = arr(gp.row_offsets, gp.num_row_cones + 1, np.uint64)
        qp.soc_rhs = arr(gp.soc_rhs, m, float)
    return qp


def random_qp(seed=0, n=2000, m_eq=600, m_in=400, a_row_nnz=10, heavy_rows=0, heavy_nnz=0,
              b_rows=500, b_row_nnz=5, box=10.0) -> QP:
    g = gen_lib()
    gp = _GenProblem()
    rc = g.qpgen_random_qp(seed, n, m_eq, m_in, a_row_nnz, heavy_rows, heavy_nnz, b_rows, b_row_nnz, box,
                           C.byref(gp))
    if rc != 0:
        raise RuntimeError("qpgen_random_qp failed")
    try:
        return _take(gp)
    finally:
        g.qpgen_free(C.byref(gp))


def soc_qp(seed=0, n_box=100, cones=10, cone_dim=10, m=50, a_row_nnz=10, box=5.0) -> QP:
    g = gen_lib()
    gp = _GenProblem()
    if g.qpgen_soc_qp(seed, n_box, cones, cone_dim, m, a_row_nnz, box, C.byref(gp)) != 0:
        raise RuntimeError("qpgen_soc_qp failed")
    try:
        return _take(gp)
    finally:
        g.qpgen_free(C.byref(gp))


# --------------------------------------------------------------------------- oracle

class OrParams(C.Structure):
    _fields_ = [("stoptol", C.c_double), ("max_iter", C.c_size_t), ("time_limit", C.c_double),
                ("check_iter", C.c_size_t), ("print_frequency", C.c_int), ("use_scaling", C.c_int),
                ("lambda_A_given", C.c_double), ("lambda_Q_given", C.c_double), ("gram_mode", C.c_int)]


class OrResults(C.Structure):
    _fields_ = [("status", C.c_int), ("iter", C.c_size_t), ("iter_4", C.c_size_t), ("iter_6", C.c_size_t),
                ("time", C.c_double), ("time_4", C.c_double), ("time_6", C.c_double),
                ("primal_obj", C.c_double), ("residuals", C.c_double), ("gap", C.c_double),
                ("err_rp", C.c_double), ("err_rd", C.c_double), ("dual_obj", C.c_double)]


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        path = ROOT / "oracle" / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(path))
        L.or_params_defaults.argtypes = [C.POINTER(OrParams)]
        L.or_session_create.argtypes = ([C.c_size_t, C.c_size_t] + [C.c_size_t, c_size_p, c_size_p, c_double_p] * 2
                                        + [c_double_p] * 5 + [C.c_double, C.c_size_t, c_size_p, C.c_size_t,
                                                              c_size_p, c_double_p, C.POINTER(OrParams),
                                                              C.POINTER(C.c_void_p)])
        L.or_session_destroy.argtypes = [C.c_void_p]
        L.or_session_advance.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_int)]
        L.or_session_solve.argtypes = [C.c_void_p, C.POINTER(OrResults)] + [c_double_p] * 4
        L.or_session_results.argtypes = [C.c_void_p, C.POINTER(OrResults)] + [c_double_p] * 4
        L.or_session_n.argtypes = [C.c_void_p]
        L.or_session_n.restype = C.c_size_t
        L.or_session_m.argtypes = [C.c_void_p]
        L.or_session_m.restype = C.c_size_t
        L.or_session_read.argtypes = [C.c_void_p, C.c_char_p, c_double_p, C.c_size_t]
        L.or_session_read.restype = C.c_size_t
        L.or_session_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.or_session_scalar.restype = C.c_double
        L.or_session_matrix.argtypes = [C.c_void_p, C.c_char_p, c_size_p, c_size_p, c_double_p]
        L.or_session_matrix.restype = C.c_size_t
        L.or_csr_matvec.argtypes = [C.c_size_t, c_size_p, c_size_p, c_double_p, c_double_p, c_double_p]
        L.or_csr_transpose.argtypes = [C.c_size_t, C.c_size_t, c_size_p, c_size_p, c_double_p, c_size_p, c_size_p,
                                       c_double_p]
        L.or_soc_project.argtypes = [c_double_p, c_double_p, C.c_size_t]
        L.or_session_write.argtypes = [C.c_void_p, C.c_char_p, c_double_p, C.c_size_t]
        L.or_session_step.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
        L.or_session_residuals.argtypes = [C.c_void_p, C.c_size_t, c_double_p]
        L.or_session_m_norm.argtypes = [C.c_void_p]
        L.or_session_m_norm.restype = C.c_double
        L.or_session_gram_quadratic.argtypes = [C.c_void_p]
        L.or_session_gram_quadratic.restype = C.c_double
        L.or_session_gram_map.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.c_size_t]
        L.or_session_gram_map.restype = C.c_size_t
        L.or_preprocess.argtypes = ([C.c_size_t, C.c_size_t] + [C.c_size_t, c_size_p, c_size_p, c_double_p] * 2
                                    + [c_double_p, c_double_p, c_size_p, c_size_p, c_size_p, c_double_p,
                                       c_size_p, c_size_p, c_size_p, c_size_p, c_double_p])
        _oracle = L
    return _oracle


def oracle_params(**kw) -> OrParams:
    p = OrParams()
    oracle_lib().or_params_defaults(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class OracleSession:
    def __init__(self, qp: QP, **params):
        L = oracle_lib()
        self._keep = []

        def sz(a):
            arr, p = _sz(a if a is not None else np.zeros(1, np.uint64))
            self._keep.append(arr)
            return p

        def db(a):
            arr, p = _db(a if a is not None and len(a) else np.zeros(1))
            self._keep.append(arr)
            return p
        self.params = oracle_params(**params)
        h = C.c_void_p()
        st = L.or_session_create(qp.n, qp.m, len(qp.q_val), sz(qp.q_ptr), sz(qp.q_ind), db(qp.q_val),
                                 len(qp.a_val), sz(qp.a_ptr), sz(qp.a_ind), db(qp.a_val), db(qp.c), db(qp.AL),
                                 db(qp.AU), db(qp.l), db(qp.u), qp.obj_constant, qp.num_var_cones,
                                 sz(qp.var_offsets) if qp.num_var_cones else None, qp.num_row_cones,
                                 sz(qp.row_offsets) if qp.num_row_cones else None,
                                 db(qp.soc_rhs) if qp.num_row_cones else None, C.byref(self.params), C.byref(h))
        self.status = st
        self.h = h if st == 0 else None
        self.n_in, self.m_in = qp.n, qp.m

    def __del__(self):
        if getattr(self, "h", None):
            oracle_lib().or_session_destroy(self.h)
            self.h = None

    @property
    def n(self):
        return oracle_lib().or_session_n(self.h)

    @property
    def m(self):
        return oracle_lib().or_session_m(self.h)

    def advance(self, bodies=1):
        fin = C.c_int(0)
        st = oracle_lib().or_session_advance(self.h, bodies, C.byref(fin))
        return st, bool(fin.value)

    def read(self, name):
        L = oracle_lib()
        ln = L.or_session_read(self.h, name.encode(), None, 0)
        out = np.zeros(ln)
        if ln:
            L.or_session_read(self.h, name.encode(), out.ctypes.data_as(c_double_p), ln)
        return out

    def scalar(self, name):
        return oracle_lib().or_session_scalar(self.h, name.encode())

    def matrix(self, name):
        L = oracle_lib()
        rows = {"Q": self.n, "A": self.m, "AT": self.n}[name]
        nnz = L.or_session_matrix(self.h, name.encode(), None, None, None)
        ptr = np.zeros(rows + 1, np.uint64)
        ind = np.zeros(max(nnz, 1), np.uint64)
        val = np.zeros(max(nnz, 1))
        L.or_session_matrix(self.h, name.encode(), ptr.ctypes.data_as(c_size_p), ind.ctypes.data_as(c_size_p),
                            val.ctypes.data_as(c_double_p))
        return ptr, ind[:nnz], val[:nnz]

    def results(self):
        r = OrResults()
        n, m = self.n, self.m
        x, y, z, w = np.zeros(max(n, 1)), np.zeros(max(m, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
        oracle_lib().or_session_results(self.h, C.byref(r), x.ctypes.data_as(c_double_p),
                                        y.ctypes.data_as(c_double_p), z.ctypes.data_as(c_double_p),
                                        w.ctypes.data_as(c_double_p))
        return r, x[:n], y[:m], z[:n], w[:n]

    def solve(self):
        self.advance((1 << 63))
        return self.results()

    def write(self, name, values):
        arr = np.ascontiguousarray(values, dtype=np.float64)
        return oracle_lib().or_session_write(self.h, name.encode(), arr.ctypes.data_as(c_double_p), arr.size)

    def step(self, inner=0, to_check=1):
        return oracle_lib().or_session_step(self.h, inner, to_check)

    def residuals(self, iter_):
        out = np.zeros(6)
        oracle_lib().or_session_residuals(self.h, iter_, out.ctypes.data_as(c_double_p))
        return dict(zip(["pobj", "dobj", "err_rd", "err_rp", "gap", "kkt"], out))


# --------------------------------------------------------------------------- reference (compiled from source)

class _CSR(C.Structure):
    """HPRQPMatrixCSR (the reference's include/shared/types.h)."""
    _fields_ = [("nrow", C.c_size_t), ("ncol", C.c_size_t), ("nnz", C.c_size_t), ("row_ptr", c_size_p),
                ("col_ind", c_size_p), ("values", c_double_p), ("owns_data", C.c_int)]


class _CModel(C.Structure):
    """HPRQPModel (the reference's include/shared/types.h:50-66)."""
    _fields_ = [("Q", _CSR), ("A", _CSR), ("AT", _CSR), ("c", c_double_p), ("AL", c_double_p),
                ("AU", c_double_p), ("l", c_double_p), ("u", c_double_p), ("obj_constant", C.c_double),
                ("owns_data", C.c_int), ("num_var_cones", C.c_size_t), ("var_cone_offsets", c_size_p),
                ("num_row_cones", C.c_size_t), ("row_cone_offsets", c_size_p), ("soc_rhs", c_double_p)]


def qp_of_cmodel(m) -> QP:
    """Copy a built HPRQPModel (any library's) into a QP, cones included."""
    def arr(p, k, dt):
        return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True) if k else np.zeros(0, dt)

    def csr(M):
        return arr(M.row_ptr, M.nrow + 1, np.uint64), arr(M.col_ind, M.nnz, np.uint64), arr(M.values, M.nnz, float)
    n, mm = m.Q.ncol, m.A.nrow
    qrp, qci, qv = csr(m.Q)
    arp, aci, av = csr(m.A)
    qp = QP(n, mm, qrp, qci, qv, arp, aci, av, arr(m.c, n, float), arr(m.AL, mm, float), arr(m.AU, mm, float),
            arr(m.l, n, float), arr(m.u, n, float), m.obj_constant)
    if m.num_var_cones:
        qp.var_offsets = arr(m.var_cone_offsets, m.num_var_cones + 1, np.uint64)
    if m.num_row_cones:
        qp.row_offsets = arr(m.row_cone_offsets, m.num_row_cones + 1, np.uint64)
        qp.soc_rhs = arr(m.soc_rhs, mm, float)
    return qp


def ref_model_from_mps(path):
    """The reference loader's model (build_from_mps of oracle/_ref), or None
    when it rejects the file."""
    L = ref_lib()
    m = _CModel()
    st = L.build_from_mps(str(path).encode(), C.byref(m))
    try:
        return qp_of_cmodel(m) if st == 0 else None
    finally:
        L.model_free(C.byref(m))


def assert_same_model(a: QP, b: QP):
    """Bitwise equality of two built models, cones included."""
    assert (a.n, a.m, a.num_var_cones, a.num_row_cones) == (b.n, b.m, b.num_var_cones, b.num_row_cones)
    assert a.obj_constant == b.obj_constant or (math.isnan(a.obj_constant) and math.isnan(b.obj_constant))
    for f in ("q_ptr", "q_ind", "q_val", "a_ptr", "a_ind", "a_val", "c", "AL", "AU", "l", "u", "var_offsets",
              "row_offsets", "soc_rhs"):
        x, y = getattr(a, f), getattr(b, f)
        if x is None or y is None:
            assert x is None and y is None, f
            continue
        np.testing.assert_array_equal(x, y, err_msg=f)
        assert np.array_equal(np.signbit(x), np.signbit(y)) if x.dtype.kind == "f" else True, f


REF_PATH = ROOT / "oracle" / "_ref" / "libhprqp_ref.so"
_ref = None


def ref_available():
    return REF_PATH.exists()


def ref_lib():
    global _ref
    if _ref is None:
        _ref = binding_lib(REF_PATH)
        _ref.build_from_mps.argtypes = [C.c_char_p, C.POINTER(_CModel)]
        _ref.model_free.argtypes = [C.POINTER(_CModel)]
    return _ref


def binding_lib(path):
    """A fresh ctypes handle on `path` (the reference library or this
    package's libhprqp_hip.so) with the argument types of the
    include/hprqp_binding.h functions set; independent of other handles on
    the same library."""
    L = C.CDLL(str(path))
    vp = C.c_void_p
    L.hprqp_model_create.argtypes = [C.POINTER(vp)]
    L.hprqp_model_destroy.argtypes = [vp]
    L.hprqp_model_build_from_qabc_csr.argtypes = ([vp] + [C.c_size_t] * 3 + [c_size_p, c_size_p, c_double_p]
                                                  + [C.c_size_t] * 3 + [c_size_p, c_size_p, c_double_p]
                                                  + [c_double_p] * 5 + [C.c_double])
    L.hprqp_model_num_variables.argtypes = [vp]
    L.hprqp_model_num_variables.restype = C.c_size_t
    L.hprqp_model_num_constraints.argtypes = [vp]
    L.hprqp_model_num_constraints.restype = C.c_size_t
    L.hprqp_params_create.argtypes = [C.POINTER(vp)]
    L.hprqp_params_destroy.argtypes = [vp]
    for nm, t in [("stoptol", C.c_double), ("max_iter", C.c_size_t), ("use_gpu", C.c_int),
                  ("time_limit", C.c_double), ("check_iter", C.c_size_t), ("print_frequency", C.c_int),
                  ("use_scaling", C.c_int), ("verbose", C.c_int)]:
        getattr(L, "hprqp_params_set_" + nm).argtypes = [vp, t]
    L.hprqp_results_create.argtypes = [C.POINTER(vp)]
    L.hprqp_results_destroy.argtypes = [vp]
    L.hprqp_optimize.argtypes = [vp, vp, vp]
    L.hprqp_model_read_from_mps.argtypes = [vp, C.c_char_p]
    L.hprqp_status_name.argtypes = [C.c_int]
    L.hprqp_status_name.restype = C.c_char_p
    L.hprqp_params_set_defaults.argtypes = [vp]
    for nm in ("device_number", "spmv_mode_q", "spmv_mode_a"):
        getattr(L, "hprqp_params_set_" + nm).argtypes = [vp, C.c_int]
    for nm in ["status", "num_variables", "num_constraints", "iter", "iter_4", "iter_6"]:
        f = getattr(L, "hprqp_results_" + nm)
        f.argtypes = [vp]
        f.restype = C.c_size_t if nm != "status" else C.c_int
    for nm in ["time", "time_4", "time_6", "objective", "residual", "gap"]:
        f = getattr(L, "hprqp_results_" + nm)
        f.argtypes = [vp]
        f.restype = C.c_double
    for nm in "xyzw":
        f = getattr(L, f"hprqp_results_{nm}_data")
        f.argtypes = [vp]
        f.restype = c_double_p
    return L


def binding_solve(L, qp: QP | None, mps_path=None, **params):
    """Solve through the hprqp_binding.h API of library L (reference or
    product), the model given as arrays (qp) or as an MPS file (mps_path)."""
    vp = C.c_void_p
    model, prm, res = vp(), vp(), vp()
    L.hprqp_model_create(C.byref(model))
    L.hprqp_params_create(C.byref(prm))
    L.hprqp_results_create(C.byref(res))
    try:
        keep = []

        def sz(a):
            arr, p = _sz(a if len(a) else np.zeros(1, np.uint64))
            keep.append(arr)
            return p

        def db(a):
            arr, p = _db(a if len(a) else np.zeros(1))
            keep.append(arr)
            return p
        if mps_path is not None:
            L.hprqp_model_read_from_mps.argtypes = [vp, C.c_char_p]
            st = L.hprqp_model_read_from_mps(model, str(mps_path).encode())
        else:
            st = L.hprqp_model_build_from_qabc_csr(model, qp.n, qp.n, len(qp.q_val), sz(qp.q_ptr), sz(qp.q_ind),
                                                   db(qp.q_val), qp.m, qp.n, len(qp.a_val), sz(qp.a_ptr),
                                                   sz(qp.a_ind), db(qp.a_val), db(qp.c), db(qp.AL), db(qp.AU),
                                                   db(qp.l), db(qp.u), qp.obj_constant)
        if st != 0:
            return {"status": STATUS[st], "build_failed": True}
        params.setdefault("verbose", 0)
        for k, v in params.items():
            getattr(L, "hprqp_params_set_" + k)(prm, v)
        st = L.hprqp_optimize(model, prm, res)
        n = L.hprqp_results_num_variables(res)
        m = L.hprqp_results_num_constraints(res)

        def vec(nm, ln):
            p = getattr(L, f"hprqp_results_{nm}_data")(res)
            return np.ctypeslib.as_array(p, shape=(ln,)).copy() if (ln and p) else np.zeros(0)
        out = {"rc": STATUS.get(st, st), "status": STATUS.get(L.hprqp_results_status(res)),
               "iter": L.hprqp_results_iter(res), "iter_4": L.hprqp_results_iter_4(res),
               "iter_6": L.hprqp_results_iter_6(res), "time": L.hprqp_results_time(res),
               "objective": L.hprqp_results_objective(res), "residual": L.hprqp_results_residual(res),
               "gap": L.hprqp_results_gap(res), "x": vec("x", n), "y": vec("y", m), "z": vec("z", n),
               "w": vec("w", n), "n": n, "m": m}
        return out
    finally:
        L.hprqp_results_destroy(res)
        L.hprqp_params_destroy(prm)
        L.hprqp_model_destroy(model)


# --------------------------------------------------------------------------- cone fixtures

def _csr_dense(M):
    M = np.asarray(M, dtype=float)
    rows, cols = np.nonzero(M)
    ptr = np.zeros(M.shape[0] + 1, np.uint64)
    np.add.at(ptr, rows + 1, 1)
    return np.cumsum(ptr).astype(np.uint64), cols.astype(np.uint64), M[rows, cols]


def qp_dense_soc(Qd, Ad, c, AL, AU, l, u, var_offsets=None, row_offsets=None, rhs=None):
    Qd = np.atleast_2d(np.asarray(Qd, float))
    n = Qd.shape[1]
    Ad = np.asarray(Ad, float).reshape(-1, n)
    qp_, qi, qv = _csr_dense(Qd)
    ap, ai, av = _csr_dense(Ad)
    qp = QP(n, Ad.shape[0], qp_, qi, qv, ap, ai, av, np.asarray(c, float), np.asarray(AL, float),
            np.asarray(AU, float), np.asarray(l, float), np.asarray(u, float))
    if var_offsets is not None:
        qp.var_offsets = np.asarray(var_offsets, np.uint64)
    if row_offsets is not None:
        qp.row_offsets = np.asarray(row_offsets, np.uint64)
        qp.soc_rhs = np.asarray(rhs, float)
    return qp


def soc_mixed_model(q_kind):
    """tests/gpu/test_gpu_soc.cu:115-150: 8 variables (box prefix 2, cones
    [2,5) and [5,8)), 7 rows (two linear, an empty one, cones [2,4), [4,7))
    built around a known primal-dual point x = (0.25,-0.5,5,3,4,2,0,0)."""
    n, m = 8, 7
    inf = math.inf
    a = np.zeros((m, n))
    for i in range(m):
        for j in range(n):
            a[i, j] = 0.0 if i == 3 else (((i * 3 + j * 5 + 1) % 11) - 5) / 7.0
    q = np.zeros((n, n))
    for j in range(n):
        if q_kind:
            q[j, j] = 1.0 + j / 8.0
        if q_kind == 2 and j + 1 < n:
            q[j, j] += 0.25
            q[j + 1, j + 1] += 0.25
            q[j, j + 1] = q[j + 1, j] = -0.25
    x = np.array([0.25, -0.5, 5, 3, 4, 2, 0, 0])
    y = np.array([0.2, 0.3, 2, -2, 0, 0, 0])
    z = np.array([0.4, 0, 1, -0.6, -0.8, 0, 0, 0])
    ax, aty, qx = a @ x, a.T @ y, q @ x
    c = aty + z - qx
    lower = np.full(n, -inf)
    upper = np.full(n, inf)
    lower[0], upper[0], lower[1], upper[1] = x[0], 1.0, -1.0, 1.0
    al = np.full(m, -inf)
    au = np.full(m, inf)
    al[0] = au[0] = ax[0]
    al[1] = ax[1]
    slack = np.array([ax[0], ax[1], 1, 1, 3, 0, 0])
    rhs = np.zeros(m)
    rhs[2:] = ax[2:] - slack[2:]
    return qp_dense_soc(q, a, c, al, au, lower, upper, [2, 5, 8], [2, 4, 7], rhs)


def soc_simple_models():
    """tests/gpu/test_gpu_soc.cu:234-273: name -> (qp, expected x or None)."""
    inf = math.inf
    out = {}
    out["proj_cone"] = (qp_dense_soc(np.eye(3), np.zeros((0, 3)), [0, -3, -4], [], [], [-inf] * 3, [inf] * 3,
                                     var_offsets=[0, 3]), [2.5, 1.5, 2.0])
    out["affine_cone"] = (qp_dense_soc([[0.0]], [[1], [0], [0]], [1], [-inf] * 3, [inf] * 3, [-inf], [inf],
                                       row_offsets=[0, 3], rhs=[2, -3, -4]), [7.0])
    out["constant_cone"] = (qp_dense_soc([[1.0]], np.zeros((2, 1)), [-1], [-inf] * 2, [inf] * 2, [-inf], [inf],
                                         row_offsets=[0, 2], rhs=[-3.0, 0.0]), [1.0])
    out["constant_cone_infeasible"] = (qp_dense_soc([[1.0]], np.zeros((2, 1)), [-1], [-inf] * 2, [inf] * 2,
                                                    [-inf], [inf], row_offsets=[0, 2], rhs=[1.0, 0.0]), None)
    return out


def gram_fixture_qp(kind):
    """The analytic QPs of the reference's Gram contract test
    (tests/polishing/fixtures.py:6-17, used by tests/polishing/test_gpu.py:14-66):
    two variables in [-1, 1]^2, Q = [[2, .3], [.3, 1]] and c = -Q x* with
    x* = (0.25, 0.5), under five row sets -- 'affine' (one equality row),
    'mixed' (an equality between two interval rows), 'dependent' (two
    proportional equality rows, so G is singular), 'no_equalities' and 'box'
    (no rows). Returns (qp, optimal objective)."""
    Qd = np.array([[2.0, 0.3], [0.3, 1.0]])
    opt = np.array([0.25, 0.5])
    c = -Qd @ opt
    if kind == "affine":
        Ad = np.array([[1.0, 2.0]])
        lo = hi = Ad @ opt
    elif kind == "mixed":
        Ad = np.array([[0.0, 1.0], [1.0, 2.0], [1.0, 0.0]])
        lo, hi = np.array([0.0, 1.25, -1.0]), np.array([1.0, 1.25, 1.0])
    elif kind == "dependent":
        Ad = np.array([[1.0, 2.0], [2.0, 4.0]])
        lo = hi = Ad @ opt
    elif kind == "no_equalities":
        Ad = np.eye(2)
        lo, hi = np.zeros(2), np.ones(2)
    elif kind == "box":
        Ad = np.zeros((0, 2))
        lo = hi = np.zeros(0)
    else:
        raise ValueError(kind)
    qp = qp_from_dense(Qd, Ad, c, lo, hi, [-1.0, -1.0], [1.0, 1.0])
    return qp, float(0.5 * opt @ Qd @ opt + c @ opt)


def gram_fixture_spatial(g):
    """The boundary-control model on a g x g grid that the reference's AUTO
    rule recognises (tests/polishing/fixtures.py:19-29, src/shared/gram_policy.cpp):
    variables [controls u (k = 4g) | states y (g^2) | slacks v (k)]; g^2
    homogeneous equality rows K y - B u = 0 with K the five-point Laplacian
    (4 on the diagonal, -1 to each grid neighbour) and B placing control j on
    boundary state br[j]; then one 3-dimensional row cone per control,
    (1 - C_j, -1 - C_j, 2 u_j) with C_j = u_j - y_br[j] - v_j. Q = diag(0 on u,
    h^2 on y, h/100 on v), c = -h^2 d on y for the target
    d = 3 - 2 |p - (1/2, 1/2)|^2 at grid point p, h = 1/(g+1)."""
    h = 1.0 / (g + 1)
    e, k = g * g, 4 * g
    n, m = 2 * k + e, e + 3 * k
    A = np.zeros((m, n))
    for r in range(g):
        for col in range(g):
            i = r * g + col
            A[i, k + i] = 4.0
            for rr, cc in ((r - 1, col), (r + 1, col), (r, col - 1), (r, col + 1)):
                if 0 <= rr < g and 0 <= cc < g:
                    A[i, k + rr * g + cc] = -1.0
    br = np.concatenate([np.arange(g), np.arange(e - g, e), np.arange(0, e, g), np.arange(g - 1, e, g)])
    for j in range(k):
        A[br[j], j] -= 1.0                  # - B u on the state rows
        t = e + 3 * j
        # rows t, t+1: -C_j = -u_j + y_br[j] + v_j ; row t+2: 2 u_j
        for rr in (t, t + 1):
            A[rr, j] = -1.0
            A[rr, k + br[j]] = 1.0
            A[rr, k + e + j] = 1.0
        A[t + 2, j] = 2.0
    v = np.arange(1, g + 1) * h
    d = (3.0 - 2.0 * ((v[:, None] - 0.5) ** 2 + (v[None, :] - 0.5) ** 2)).ravel()
    qd = np.concatenate([np.zeros(k), np.full(e, h * h), np.full(k, 0.01 * h)])
    c = np.concatenate([np.zeros(k), -h * h * d, np.zeros(k)])
    l = np.concatenate([np.zeros(k + e), np.full(k, 3.7)])
    u = np.concatenate([np.full(k + e, 2.071), np.full(k, 4.5)])
    AL = np.concatenate([np.zeros(e), np.full(3 * k, -math.inf)])
    AU = np.concatenate([np.zeros(e), np.full(3 * k, math.inf)])
    rhs = np.concatenate([np.zeros(e), np.tile([-1.0, 1.0, 0.0], k)])
    qp = qp_dense_soc(np.diag(qd), A, c, AL, AU, l, u, row_offsets=np.arange(e, m + 1, 3), rhs=rhs)
    qp.obj_constant = float(0.5 * h * h * (d @ d))
    return qp


def cone_project(v):
    v = np.asarray(v, float)
    t, x = v[0], v[1:]
    nx = np.linalg.norm(x)
    if nx <= t:
        return v.copy()
    if nx <= -t:
        return np.zeros_like(v)
    return np.concatenate([[(nx + t) / 2], (nx + t) / (2 * nx) * x])


def soc_certificate(qp: QP, x, y, z):
    """Independent original-units optimality certificate
    (tests/gpu/test_gpu_soc.cu:152-196): primal/dual cone residuals,
    stationarity and the relative gap using support values."""
    import scipy.sparse as sp
    n, m = qp.n, qp.m
    Qm = sp.csr_matrix((qp.q_val, qp.q_ind.astype(np.int64), qp.q_ptr.astype(np.int64)), shape=(n, n))
    Am = sp.csr_matrix((qp.a_val, qp.a_ind.astype(np.int64), qp.a_ptr.astype(np.int64)), shape=(m, n))
    lin_n = int(qp.var_offsets[0]) if qp.num_var_cones else n
    lin_m = int(qp.row_offsets[0]) if qp.num_row_cones else m
    qx, ax, aty = 0.5 * (Qm @ x + Qm.T @ x), Am @ x, Am.T @ y
    rd = qx + qp.c - aty - z
    primal = dual = support = 0.0

    def bounds_part(v, lo, hi, dual_v):
        # primal bound violation, support-function value and the dual
        # residual of the multipliers that no finite bound supports
        pos = (dual_v > 0) & np.isfinite(lo)
        neg = ~pos & (dual_v < 0) & np.isfinite(hi)
        other = ~pos & ~neg
        viol = float(np.max(np.maximum(lo - v, v - hi))) if len(v) else 0.0
        sup = float(np.sum(dual_v[pos] * lo[pos]) + np.sum(dual_v[neg] * hi[neg]))
        res = float(np.max(np.abs(dual_v[other]))) if other.any() else 0.0
        return viol, sup, res
    for vv, lo, hi, dv in ((x[:lin_n], qp.l[:lin_n], qp.u[:lin_n], z[:lin_n]),
                           (ax[:lin_m], qp.AL[:lin_m], qp.AU[:lin_m], y[:lin_m])):
        viol, sup, res = bounds_part(vv, lo, hi, dv)
        primal, dual, support = max(primal, viol), max(dual, res), support + sup
    for k in range(qp.num_var_cones):
        b, e = int(qp.var_offsets[k]), int(qp.var_offsets[k + 1])
        primal = max(primal, np.max(np.abs(cone_project(x[b:e]) - x[b:e])))
        dual = max(dual, np.max(np.abs(cone_project(z[b:e]) - z[b:e])))
    slack = ax.copy()
    for i in range(lin_m, m):
        slack[i] -= qp.soc_rhs[i]
        support += y[i] * qp.soc_rhs[i]
    for k in range(qp.num_row_cones):
        b, e = int(qp.row_offsets[k]), int(qp.row_offsets[k + 1])
        primal = max(primal, np.max(np.abs(cone_project(slack[b:e]) - slack[b:e])))
        dual = max(dual, np.max(np.abs(cone_project(y[b:e]) - y[b:e])))
    pobj = 0.5 * x @ qx + qp.c @ x + qp.obj_constant
    dobj = -0.5 * x @ qx + support + qp.obj_constant
    gap = abs(pobj - dobj) / (1 + max(abs(pobj), abs(dobj)))
    inf_ = lambda a: float(np.max(np.abs(a))) if len(a) else 0.0
    stat = inf_(rd) / (1 + max(inf_(qp.c), inf_(qx), inf_(aty)))
    return dict(primal=primal, dual=dual, stationarity=stat, gap=gap)


# The benchmark workloads (generator arguments), shared by bench.py and the
# full-size tests.
# plain_n2k: n = 2,000, m = 1,000 (600 equality + 400 two-sided rows), A at
# 10 nnz/row, Q = B'DB + 1e-2 I with B 500 x n at 5 nnz/row, box [-10, 10].
PLAIN_N2K = dict(seed=0, n=2_000, m_eq=600, m_in=400, a_row_nnz=10, heavy_rows=0, heavy_nnz=0, b_rows=500,
                 b_row_nnz=5, box=10.0)
# plain_n1M: n = 1e6, m = 5e5 (300k + 200k), A at 20 nnz/row with 5,000 rows
# of 2,000 nnz, B 200,000 x n at 8 nnz/row.
PLAIN_N1M = dict(seed=0, n=1_000_000, m_eq=300_000, m_in=200_000, a_row_nnz=20, heavy_rows=5_000, heavy_nnz=2_000,
                 b_rows=200_000, b_row_nnz=8, box=10.0)
# soc_n220k: 20,000 cones of dimension 10 after a 20,000-variable box
# prefix [-5, 5] (n = 220,000), 100,000 equality rows at 40 nnz/row, Q
# diagonal U(0, 1).
SOC_N220K = dict(seed=0, n_box=20_000, cones=20_000, cone_dim=10, m=100_000, a_row_nnz=40, box=5.0)


def batched_instances(count=4096, seed=0, n_lo=50, n_hi=2000):
    """The batched workload: n ~ U{n_lo..n_hi}, m = round(n/2) (60% equality),
    A at min(10, n) nnz/row, Q = B'DB + 1e-2 I with B (n/4) x n at 4 nnz/row,
    box [-10, 10]; instance k uses generator seed `seed * 1_000_003 + k`."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        m = int(round(n / 2))
        m_eq = int(round(0.6 * m))
        out.append(random_qp(seed=seed * 1_000_003 + k, n=n, m_eq=m_eq, m_in=m - m_eq, a_row_nnz=min(10, n),
                             heavy_rows=0, heavy_nnz=0, b_rows=max(1, n // 4), b_row_nnz=4, box=10.0))
    return out


def layer_instances(count, n=256, m_eq=64, m_in=256, seed=0):
    """Dense QP-layer workload: Q = M M'/n + 0.1 I (M n x n, N(0,1)),
    A dense N(0, 1/n), x0 ~ U(-1,1), c ~ N(0,1); equality rows b = A_eq x0,
    two-sided rows A_in x0 -/+ U(0,1); variables free."""
    rng = np.random.default_rng(seed)
    m = m_eq + m_in
    Q = np.empty((count, n, n))
    A = rng.normal(0.0, 1.0 / math.sqrt(n), size=(count, m, n))
    c = rng.normal(size=(count, n))
    AL = np.empty((count, m))
    AU = np.empty((count, m))
    for k in range(count):
        M = rng.normal(size=(n, n))
        Q[k] = M @ M.T / n + 0.1 * np.eye(n)
        x0 = rng.uniform(-1, 1, size=n)
        ax = A[k] @ x0
        AL[k, :m_eq] = AU[k, :m_eq] = ax[:m_eq]
        AL[k, m_eq:] = ax[m_eq:] - rng.uniform(0, 1, size=m_in)
        AU[k, m_eq:] = ax[m_eq:] + rng.uniform(0, 1, size=m_in)
    l = np.full((count, n), -math.inf)
    u = np.full((count, n), math.inf)
    return Q, c, A, AL, AU, l, u


def kkt_active_solution(Q, c, A, AL, AU, l, u, y, z, tol=0.0):
    """Exact solution of the equality-constrained QP on the active set read
    from the duals (row i active if y_i != 0 or AL_i == AU_i; variable j if
    z_j != 0 or l_j == u_j), by a dense KKT solve. Returns x, the KKT matrix,
    the active row / variable index lists and their bound values."""
    n = len(c)
    Qs = 0.5 * (Q + Q.T)
    rows = [i for i in range(len(AL)) if abs(y[i]) > tol or AL[i] == AU[i]]
    vars_ = [j for j in range(n) if abs(z[j]) > tol or l[j] == u[j]]
    G = np.vstack([A[rows]] + [np.eye(n)[vars_]]) if (rows or vars_) else np.zeros((0, n))
    h = np.concatenate([[AL[i] if (AL[i] == AU[i] or y[i] > 0) else AU[i] for i in rows],
                        [l[j] if (l[j] == u[j] or z[j] > 0) else u[j] for j in vars_]])
    k = G.shape[0]
    K = np.zeros((n + k, n + k))
    K[:n, :n] = Qs
    K[:n, n:] = G.T
    K[n:, :n] = G
    sol = np.linalg.solve(K, np.concatenate([-c, h]))
    return sol[:n], K, rows, vars_


# ------------------------------------------------------------ device buffers
# Plain HIP runtime calls (libamdhip64) for the kernel-level C-ABI entry
# points, which take device pointers.
_hip = None


def hip_runtime():
    global _hip
    if _hip is None:
        # the HIP runtime the library is bound to (PyTorch's when installed,
        # see hprqp_amd._load): opened by soname, which resolves to the copy
        # already mapped instead of mapping a second runtime
        import hprqp_amd  # noqa: F401
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


class DeviceArray:
    """A device copy of a numpy array (freed on close)."""

    H2D, D2H = 1, 2

    def __init__(self, a=None, *, shape=None, dtype=None):
        self.host = np.ascontiguousarray(a) if a is not None else np.zeros(shape, dtype=dtype)
        self.ptr = C.c_void_p()
        nbytes = max(self.host.nbytes, 8)
        if hip_runtime().hipMalloc(C.byref(self.ptr), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        if a is not None and self.host.nbytes:
            if hip_runtime().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, self.H2D) != 0:
                raise RuntimeError("hipMemcpy H2D failed")

    def get(self):
        out = np.empty_like(self.host)
        if out.nbytes and hip_runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def close(self):
        if self.ptr:
            hip_runtime().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def csr_with_row_lengths(lengths, ncol, seed=0):
    """CSR with the given row lengths, distinct ascending random columns per
    row, values N(0,1)."""
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, dtype=np.int64)
    rp = np.zeros(len(lengths) + 1, dtype=np.int64)
    rp[1:] = np.cumsum(lengths)
    ci = np.empty(rp[-1], dtype=np.int64)
    for r, L in enumerate(lengths):
        if L:
            ci[rp[r]:rp[r + 1]] = np.sort(rng.choice(ncol, size=int(L), replace=False))
    val = rng.standard_normal(rp[-1])
    return rp, ci, val


# --------------------------------------------------------------------------- MPS files

# bound cards per variable kind (MPS semantics in the comment)
_BOUND_KINDS = {
    "default": [],                               # [0, inf)
    "box": [("LO", -2.0), ("UP", 3.0)],          # [-2, 3]
    "fixed": [("FX", 0.5)],                      # [0.5, 0.5]
    "free": [("FR", None)],                      # (-inf, inf)
    "minus": [("MI", None), ("UP", 4.0)],        # (-inf, 4]
    "up_neg": [("UP", -1.5)],                    # (-inf, -1.5]
    "up_pos": [("UP", 2.5)],                     # [0, 2.5]
    "lo_only": [("LO", -3.0), ("PL", None)],     # [-3, inf)
}


def write_mps(qp: QP, path, seed=0, quad="QUADOBJ", extras=True, qc=None):
    """Write `qp` (its A, c, row bounds and Q) as a free-format MPS / QPS file
    that exercises the format: ranges on L, G and E rows, all continuous bound
    types (variable bounds drawn from _BOUND_KINDS, not qp.l / qp.u), an
    objective constant through the objective row's RHS, two-pair cards, D
    exponents, comments, and (extras) an ignored second objective row and
    ignored second RHS / RANGES / BOUNDS sets. quad: "QUADOBJ" (upper
    triangle) or "QMATRIX" (full matrix). qc: {row: [(column, h), ...]}
    written as QCMATRIX blocks (diagonal cards x_j x_j h, in the given order,
    repeats allowed)."""
    rng = np.random.default_rng(seed)
    n, m = qp.n, qp.m
    A = qp.dense_A()
    Qd = qp.dense_Q()
    Qd = 0.5 * (Qd + Qd.T)
    kinds, rhs, ranges = [], {}, {}
    for i in range(m):
        lo, hi = qp.AL[i], qp.AU[i]
        if lo == hi:
            kinds.append("E")
            rhs[i] = lo
            if i % 7 == 3:
                ranges[i] = float(rng.uniform(-1, 1))
        elif math.isinf(lo):
            kinds.append("L")
            rhs[i] = hi
        elif math.isinf(hi):
            kinds.append("G")
            rhs[i] = lo
        else:
            kinds.append("L" if i % 2 else "G")
            rhs[i] = hi if i % 2 else lo
            ranges[i] = (hi - lo) * (1 if i % 4 < 2 else -1)
    num = lambda v: repr(float(v)).replace("e", "D") if abs(v) > 1e3 or (0 < abs(v) < 1e-3) else repr(float(v))
    L = ["* generated test model", "NAME          GENQP", "ROWS", " N  COST"]
    if extras:
        L.append(" N  COST2")
    L += [f" {k}  R{i}" for i, k in enumerate(kinds)]
    L.append("COLUMNS")
    for j in range(n):
        pairs = [("COST", qp.c[j])] if qp.c[j] != 0.0 else []
        if extras and j % 4 == 0:
            pairs.append(("COST2", 1.0))
        pairs += [(f"R{i}", A[i, j]) for i in np.nonzero(A[:, j])[0]]
        if not pairs:
            pairs = [("COST", 0.0)]
        for k in range(0, len(pairs), 2):
            L.append(f"    X{j}  " + "  ".join(f"{r}  {num(v)}" for r, v in pairs[k:k + 2]))
        if extras and j == n // 2:
            L += ["", "* a comment between cards"]
    L.append("RHS")
    items = [(f"R{i}", v) for i, v in rhs.items() if v != 0.0] + [("COST", -1.25)]
    for k in range(0, len(items), 2):
        L.append("    RHS  " + "  ".join(f"{r}  {num(v)}" for r, v in items[k:k + 2]))
    if extras:
        L.append("    RHS2  R0  123.0")
    if ranges:
        L.append("RANGES")
        L += [f"    RNG  R{i}  {num(r)}" for i, r in ranges.items()]
        if extras:
            L.append("    RNG2  R0  5.0")
    L.append("BOUNDS")
    names = sorted(_BOUND_KINDS)
    for j in range(n):
        for kind, val in _BOUND_KINDS[names[j % len(names)]]:
            L.append(f" {kind} BND  X{j}" + ("" if val is None else f"  {num(val)}"))
    if extras:
        L.append(" UP BND2  X0  -7.0")
    L.append(quad)
    for i in range(n):
        for j in range(n):
            if Qd[i, j] != 0.0 and (quad == "QMATRIX" or j >= i):
                L.append(f"    X{i}  X{j}  {num(Qd[i, j])}")
    for i, cards in (qc or {}).items():
        L.append(f"QCMATRIX   R{i}")
        L += [f"    X{j}  X{j}  {num(h)}" for j, h in cards]
    L.append("ENDATA")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")
