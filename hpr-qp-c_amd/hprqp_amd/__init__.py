"""Python mirror of the reference's ctypes wrapper (python/src/hprqp/__init__.py)
over this repository's MI355X library, libhprqp_hip.so.

Same names and meaning as the reference wrapper: ``Model``, ``Parameters``,
``build_from_qabc``, ``optimize``, ``Result``, ``status_name``,
``gpu_backend_available`` and the ``STATUS_*`` constants. Additions:
``build_from_qabc_soc`` (cone models, the reference's build_from_QAbc_soc),
``Session`` (step the device-resident solve loop; hprqp_hip.h) and
``solve_arrays`` (CSR arrays in, Result out).

There is no CPU fallback: importing raises if the HIP library is not built,
and solving raises if no gfx950 device is present.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

STATUS_OK = 0
STATUS_INVALID_ARGUMENT = 1
STATUS_ALLOCATION_FAILED = 2
STATUS_NOT_IMPLEMENTED = 3
STATUS_MAX_ITER = 4
STATUS_TIME_LIMIT = 5
STATUS_OPTIMAL = 6
STATUS_CONTINUE = 7

SPMV_MODE_AUTO = 0
SPMV_MODE_CUSPARSE = 1
SPMV_MODE_CUSTOMIZED = 2

PKG_DIR = Path(__file__).resolve().parents[1]
LIB_PATH = Path(os.environ.get("HPRQP_LIBRARY_PATH", PKG_DIR / "libhprqp_hip.so"))

_sz_p = C.POINTER(C.c_size_t)
_db_p = C.POINTER(C.c_double)


class MatrixCSR(C.Structure):
    """HPRQPMatrixCSR (include/shared/types.h)."""
    _fields_ = [("nrow", C.c_size_t), ("ncol", C.c_size_t), ("nnz", C.c_size_t), ("row_ptr", _sz_p),
                ("col_ind", _sz_p), ("values", _db_p), ("owns_data", C.c_int)]


class SOCSpec(C.Structure):
    _fields_ = [("num_var_cones", C.c_size_t), ("var_offsets", _sz_p), ("num_row_cones", C.c_size_t),
                ("row_offsets", _sz_p), ("rhs", _db_p)]


class CModel(C.Structure):
    _fields_ = [("Q", MatrixCSR), ("A", MatrixCSR), ("AT", MatrixCSR), ("c", _db_p), ("AL", _db_p),
                ("AU", _db_p), ("l", _db_p), ("u", _db_p), ("obj_constant", C.c_double), ("owns_data", C.c_int),
                ("num_var_cones", C.c_size_t), ("var_cone_offsets", _sz_p), ("num_row_cones", C.c_size_t),
                ("row_cone_offsets", _sz_p), ("soc_rhs", _db_p)]


class CParams(C.Structure):
    _fields_ = [("stoptol", C.c_double), ("max_iter", C.c_size_t), ("use_gpu", C.c_int),
                ("device_number", C.c_int), ("time_limit", C.c_double), ("check_iter", C.c_size_t),
                ("print_frequency", C.c_int), ("use_scaling", C.c_int), ("verbose", C.c_int),
                ("spmv_mode_q", C.c_int), ("spmv_mode_a", C.c_int), ("gram_mode", C.c_int),
                ("polishing", C.c_int), ("polish_python", C.c_char_p), ("polish_worker", C.c_char_p),
                ("polish_directory", C.c_char_p), ("polish_keep_files", C.c_int),
                ("polish_max_attempts", C.c_size_t), ("polish_episode_seconds", C.c_double),
                ("polish_total_seconds", C.c_double), ("polish_memory_bytes", C.c_size_t)]


class CResults(C.Structure):
    _fields_ = [("iter", C.c_size_t), ("iter_4", C.c_size_t), ("iter_6", C.c_size_t), ("time", C.c_double),
                ("time_4", C.c_double), ("time_6", C.c_double), ("primal_obj", C.c_double),
                ("residuals", C.c_double), ("gap", C.c_double), ("status", C.c_int), ("x", _db_p), ("y", _db_p),
                ("z", _db_p), ("w", _db_p), ("polish_attempts", C.c_size_t), ("polish_accepted", C.c_int),
                ("gram_active", C.c_int), ("total_time", C.c_double)]


def _load() -> C.CDLL:
    # PyTorch-ROCm ships its own HIP and HSA runtimes under the sonames of
    # ROCm's. Loaded after this library, torch would map a second runtime
    # into the process and find no GPU; loaded first, this library binds to
    # torch's runtime and the two share devices, streams and memory.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.hprqp_status_name.argtypes = [C.c_int]
    L.hprqp_status_name.restype = C.c_char_p
    L.gpu_backend_available.restype = C.c_int
    L.hprqp_hip_device_name.restype = C.c_char_p
    L.model_init.argtypes = [C.POINTER(CModel)]
    L.model_free.argtypes = [C.POINTER(CModel)]
    L.build_from_QAbc.argtypes = [C.POINTER(MatrixCSR), C.POINTER(MatrixCSR)] + [_db_p] * 5 + [C.c_double,
                                                                                              C.POINTER(CModel)]
    L.build_from_QAbc_soc.argtypes = ([C.POINTER(MatrixCSR), C.POINTER(MatrixCSR)] + [_db_p] * 5
                                      + [C.c_double, C.POINTER(SOCSpec), C.POINTER(CModel)])
    L.params_set_defaults.argtypes = [C.POINTER(CParams)]
    L.results_init.argtypes = [C.POINTER(CResults)]
    L.results_free.argtypes = [C.POINTER(CResults)]
    L.solve.argtypes = [C.POINTER(CModel), C.POINTER(CParams), C.POINTER(CResults)]
    L.solve_gpu.argtypes = [C.POINTER(CModel), C.POINTER(CParams), C.POINTER(CResults)]
    L.hprqp_hip_session_create.argtypes = [C.POINTER(CModel), C.POINTER(CParams), C.POINTER(vp)]
    L.hprqp_hip_session_destroy.argtypes = [vp]
    L.hprqp_hip_session_advance.argtypes = [vp, C.c_size_t, C.POINTER(C.c_int)]
    L.hprqp_hip_session_read.argtypes = [vp, C.c_char_p, _db_p, C.c_size_t]
    L.hprqp_hip_session_read.restype = C.c_size_t
    L.hprqp_hip_session_scalar.argtypes = [vp, C.c_char_p]
    L.hprqp_hip_session_scalar.restype = C.c_double
    L.hprqp_hip_session_matrix.argtypes = [vp, C.c_char_p, _sz_p, _sz_p, _db_p]
    L.hprqp_hip_session_matrix.restype = C.c_size_t
    L.hprqp_hip_session_results.argtypes = [vp, C.POINTER(CResults)]
    return L


_LIB = _load()


def lib() -> C.CDLL:
    return _LIB


def status_name(status: int) -> str:
    return _LIB.hprqp_status_name(int(status)).decode()


def gpu_backend_available() -> bool:
    return bool(_LIB.gpu_backend_available())


def device_name() -> str:
    return _LIB.hprqp_hip_device_name().decode()


@dataclass(frozen=True)
class Result:
    status_code: int
    status: str
    iter: int
    iter_4: int
    iter_6: int
    time: float
    time_4: float
    time_6: float
    objective: float
    residual: float
    gap: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    w: np.ndarray
    total_time: float = 0.0
    polish_attempts: int = 0
    polish_accepted: bool = False
    gram_active: bool = False


class Parameters:
    """HPRQPParams with the reference wrapper's keyword names."""

    def __init__(self, **kwargs: object) -> None:
        self._p = CParams()
        _LIB.params_set_defaults(C.byref(self._p))
        for key, value in kwargs.items():
            if not hasattr(self._p, key):
                raise TypeError(f"unknown parameter {key!r}")
            if isinstance(value, bool):
                value = int(value)
            elif isinstance(value, str):  # the path fields are char pointers
                value = value.encode()
            setattr(self._p, key, value)

    @property
    def c(self) -> CParams:
        return self._p


def _csr_of(matrix, nrow_hint=None):
    """(nrow, ncol, row_ptr, col_ind, values) from dense, scipy-like or CSR tuple."""
    if isinstance(matrix, tuple) and len(matrix) == 5:
        nrow, ncol, rp, ci, v = matrix
        return (int(nrow), int(ncol), np.asarray(rp, np.uint64), np.asarray(ci, np.uint64),
                np.asarray(v, np.float64))
    if hasattr(matrix, "tocsr"):
        m = matrix.tocsr()
        m.sort_indices()
        return (m.shape[0], m.shape[1], m.indptr.astype(np.uint64), m.indices.astype(np.uint64),
                m.data.astype(np.float64))
    d = np.asarray(matrix, dtype=np.float64)
    if d.ndim == 1 and d.size == 0:
        d = d.reshape(0, nrow_hint or 0)
    mask = d != 0
    rp = np.zeros(d.shape[0] + 1, np.uint64)
    np.cumsum(np.count_nonzero(mask, axis=1), out=rp[1:])
    return d.shape[0], d.shape[1], rp, np.nonzero(mask)[1].astype(np.uint64), d[mask]


class Model:
    """An owned HPRQPModel built by this library (preprocessed like the reference)."""

    def __init__(self) -> None:
        self._lib, self._byref = _LIB, C.byref  # held so that __del__ works during interpreter shutdown
        self._m = CModel()
        _LIB.model_init(C.byref(self._m))

    def __del__(self) -> None:
        if getattr(self, "_m", None) is not None:
            self._lib.model_free(self._byref(self._m))
            self._m = None

    @property
    def c(self) -> CModel:
        return self._m

    def num_variables(self) -> int:
        return int(self._m.Q.ncol)

    def num_constraints(self) -> int:
        return int(self._m.A.nrow)


def _mat(nrow, ncol, rp, ci, v, keep):
    rp = np.ascontiguousarray(rp if len(rp) else np.zeros(1), np.uint64)
    ci = np.ascontiguousarray(ci if len(ci) else np.zeros(1), np.uint64)
    v = np.ascontiguousarray(v if len(v) else np.zeros(1), np.float64)
    keep += [rp, ci, v]
    return MatrixCSR(nrow, ncol, int(rp[-1]) if nrow else 0, rp.ctypes.data_as(_sz_p), ci.ctypes.data_as(_sz_p),
                     v.ctypes.data_as(_db_p), 0)


def _dvec(a, keep):
    a = np.ascontiguousarray(a if a is not None and len(a) else np.zeros(1), np.float64)
    keep.append(a)
    return a.ctypes.data_as(_db_p)


def build_from_qabc(Q, c, A, AL, AU, l, u, obj_constant: float = 0.0, soc: dict | None = None) -> Model:
    """Same argument order as the reference wrapper's build_from_qabc."""
    c = np.asarray(c, np.float64)
    n = c.size
    qn, qm, qrp, qci, qv = _csr_of(Q)
    an, am, arp, aci, av = _csr_of(A, n)
    if an == 0:
        am = n
    keep: list = []
    Qm = _mat(qn, qm, qrp, qci, qv, keep)
    Am = _mat(an, am, arp if an else np.zeros(1, np.uint64), aci, av, keep)
    model = Model()
    args = (C.byref(Qm), C.byref(Am), _dvec(c, keep), _dvec(AL, keep), _dvec(AU, keep), _dvec(l, keep),
            _dvec(u, keep), float(obj_constant))
    if soc:
        vo = np.ascontiguousarray(soc.get("var_offsets", np.zeros(1)), np.uint64)
        ro = np.ascontiguousarray(soc.get("row_offsets", np.zeros(1)), np.uint64)
        rhs = soc.get("rhs")
        keep += [vo, ro]
        spec = SOCSpec(len(vo) - 1 if "var_offsets" in soc else 0, vo.ctypes.data_as(_sz_p),
                       len(ro) - 1 if "row_offsets" in soc else 0, ro.ctypes.data_as(_sz_p),
                       _dvec(rhs, keep) if rhs is not None else None)
        st = _LIB.build_from_QAbc_soc(*args, C.byref(spec), C.byref(model.c))
    else:
        st = _LIB.build_from_QAbc(*args, C.byref(model.c))
    if st != STATUS_OK:
        raise ValueError(f"build_from_QAbc failed: {status_name(st)}")
    return model


_LIB.build_from_QAbc_qc_factors.argtypes = ([C.POINTER(MatrixCSR), C.POINTER(MatrixCSR)] + [_db_p] * 5
                                             + [C.c_double, C.POINTER(MatrixCSR), _db_p, C.POINTER(MatrixCSR),
                                                _sz_p, C.POINTER(CModel)])
_LIB.build_from_QAbc_qc_diagonal.argtypes = ([C.POINTER(MatrixCSR), C.POINTER(MatrixCSR)] + [_db_p] * 5
                                              + [C.c_double, C.POINTER(MatrixCSR), _db_p, C.POINTER(MatrixCSR),
                                                 C.POINTER(CModel)])


def build_from_qabc_qc(Q, c, A, AL, AU, l, u, obj_constant: float, Cq, constant, R=None, offsets=None,
                       D=None) -> Model:
    """Quadratic constraints 0.5*||R_i x||^2 + C_i x + constant_i <= 0 lifted
    to row cones (build_from_QAbc_qc_factors; R's rows offsets[i] ..
    offsets[i+1]-1 form R_i) or, with D instead of R and offsets,
    0.5*sum_j D_ij x_j^2 + C_i x + constant_i <= 0
    (build_from_QAbc_qc_diagonal). Matrices as in build_from_qabc."""
    c = np.asarray(c, np.float64)
    n = c.size
    keep: list = []
    qn, qm, qrp, qci, qv = _csr_of(Q)
    an, am, arp, aci, av = _csr_of(A, n)
    if an == 0:
        am = n
    Qm = _mat(qn, qm, qrp, qci, qv, keep)
    Am = _mat(an, am, arp if an else np.zeros(1, np.uint64), aci, av, keep)
    Cm = _mat(*_csr_of(Cq, n), keep)
    args = [C.byref(Qm), C.byref(Am), _dvec(c, keep), _dvec(AL, keep), _dvec(AU, keep), _dvec(l, keep),
            _dvec(u, keep), float(obj_constant), C.byref(Cm), _dvec(constant, keep)]
    model = Model()
    if D is not None:
        Dm = _mat(*_csr_of(D, n), keep)
        st = _LIB.build_from_QAbc_qc_diagonal(*args, C.byref(Dm), C.byref(model.c))
    else:
        Rm = _mat(*_csr_of(R, n), keep)
        off = np.ascontiguousarray(offsets, np.uint64)
        keep.append(off)
        st = _LIB.build_from_QAbc_qc_factors(*args, C.byref(Rm), off.ctypes.data_as(_sz_p), C.byref(model.c))
    if st != STATUS_OK:
        raise ValueError(f"build_from_QAbc_qc failed: {status_name(st)}")
    return model


_LIB.build_from_mps.argtypes = [C.c_char_p, C.POINTER(CModel)]


def read_from_mps(path) -> Model:
    """build_from_mps on a free-format MPS / QPS file (same name as the
    reference wrapper's read_from_mps); raises ValueError on a rejected file."""
    model = Model()
    st = _LIB.build_from_mps(os.fsencode(str(path)), C.byref(model.c))
    if st != STATUS_OK:
        raise ValueError(f"build_from_mps failed: {status_name(st)}")
    return model


def build_from_mps(path) -> Model:
    return read_from_mps(path)


def _vec(p, n):
    return np.ctypeslib.as_array(p, shape=(n,)).copy() if n and p else np.zeros(0)


def _result_of(r: CResults, st: int, n: int, m: int) -> Result:
    return Result(st, status_name(st), r.iter, r.iter_4, r.iter_6, r.time, r.time_4, r.time_6, r.primal_obj,
                  r.residuals, r.gap, _vec(r.x, n), _vec(r.y, m), _vec(r.z, n), _vec(r.w, n), r.total_time,
                  int(r.polish_attempts), bool(r.polish_accepted), bool(r.gram_active))


def optimize(model: Model, params: Parameters | None = None) -> Result:
    params = params or Parameters()
    r = CResults()
    _LIB.results_init(C.byref(r))
    st = _LIB.solve(C.byref(model.c), C.byref(params.c), C.byref(r))
    try:
        return _result_of(r, st, model.num_variables(), model.num_constraints())
    finally:
        _LIB.results_free(C.byref(r))


class Session:
    """hprqp_hip_session_*: a device-resident solve advanced loop body by body."""

    def __init__(self, model: Model, params: Parameters | None = None) -> None:
        self._lib = _LIB  # held so that __del__ works during interpreter shutdown
        self.model = model
        self.params = params or Parameters(verbose=False)
        h = C.c_void_p()
        st = _LIB.hprqp_hip_session_create(C.byref(model.c), C.byref(self.params.c), C.byref(h))
        if st != STATUS_OK:
            raise RuntimeError(f"hprqp_hip_session_create failed: {status_name(st)}")
        self.h = h

    def __del__(self) -> None:
        if getattr(self, "h", None):
            self._lib.hprqp_hip_session_destroy(self.h)
            self.h = None

    def advance(self, bodies: int = 1) -> tuple[int, bool]:
        fin = C.c_int(0)
        st = _LIB.hprqp_hip_session_advance(self.h, bodies, C.byref(fin))
        return st, bool(fin.value)

    def read(self, name: str) -> np.ndarray:
        ln = _LIB.hprqp_hip_session_read(self.h, name.encode(), None, 0)
        out = np.zeros(ln)
        if ln:
            _LIB.hprqp_hip_session_read(self.h, name.encode(), out.ctypes.data_as(_db_p), ln)
        return out

    def scalar(self, name: str) -> float:
        return _LIB.hprqp_hip_session_scalar(self.h, name.encode())

    def matrix(self, name: str):
        rows = self.model.num_constraints() if name == "A" else self.model.num_variables()
        nnz = _LIB.hprqp_hip_session_matrix(self.h, name.encode(), None, None, None)
        rp = np.zeros(rows + 1, np.uint64)
        ci = np.zeros(max(nnz, 1), np.uint64)
        v = np.zeros(max(nnz, 1))
        _LIB.hprqp_hip_session_matrix(self.h, name.encode(), rp.ctypes.data_as(_sz_p), ci.ctypes.data_as(_sz_p),
                                      v.ctypes.data_as(_db_p))
        return rp, ci[:nnz], v[:nnz]

    def results(self) -> Result:
        r = CResults()
        _LIB.results_init(C.byref(r))
        st = _LIB.hprqp_hip_session_results(self.h, C.byref(r))
        try:
            return _result_of(r, r.status if st == STATUS_OK else st, self.model.num_variables(),
                              self.model.num_constraints())
        finally:
            _LIB.results_free(C.byref(r))


_LIB.hprqp_hip_batch_create.argtypes = [C.POINTER(CModel), C.c_size_t, C.POINTER(CParams), C.POINTER(C.c_void_p)]
_LIB.hprqp_hip_batch_solve.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
_LIB.hprqp_hip_batch_seconds.argtypes = [C.c_void_p]
_LIB.hprqp_hip_batch_seconds.restype = C.c_double
_LIB.hprqp_hip_batch_wave_row.argtypes = [C.c_void_p]
_LIB.hprqp_hip_batch_wave_row.restype = C.c_int
_LIB.hprqp_hip_batch_result.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(CResults)]
_LIB.hprqp_hip_batch_destroy.argtypes = [C.c_void_p]


_LIB.hprqp_hip_batch_create_dense.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t] + [_db_p] * 7 + [
    C.POINTER(CParams), C.POINTER(C.c_void_p)]
_LIB.hprqp_hip_batch_solution.argtypes = [C.c_void_p, _db_p, _db_p, _db_p, C.POINTER(C.c_int),
                                          C.POINTER(C.c_size_t), _db_p, _db_p]


class Batch:
    """hprqp_hip_batch_*: many independent QPs in one persistent launch."""

    def __init__(self, models: Sequence[Model] | None, params: Parameters | None = None) -> None:
        self._lib = _LIB  # held so that __del__ works during interpreter shutdown
        self.params = params or Parameters(verbose=False)
        self.h = None
        if models is None:  # filled in by from_dense
            return
        self.models = list(models)
        self.ns = [m.num_variables() for m in self.models]
        self.ms = [m.num_constraints() for m in self.models]
        arr = (CModel * len(self.models))(*[m.c for m in self.models])
        self._arr = arr
        h = C.c_void_p()
        st = _LIB.hprqp_hip_batch_create(arr, len(self.models), C.byref(self.params.c), C.byref(h))
        if st != STATUS_OK:
            raise RuntimeError(f"hprqp_hip_batch_create failed: {status_name(st)}")
        self.h = h

    @classmethod
    def from_dense(cls, Q, c, A, AL, AU, l, u, params: Parameters | None = None) -> "Batch":
        """hprqp_hip_batch_create_dense: count QPs of one shape from dense
        arrays Q (count, n, n), A (count, m, n), c, l, u (count, n), AL, AU
        (count, m); the same result as building every instance with
        build_from_qabc and batching the models."""
        b = cls(None, params)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Q, c, A, AL, AU, l, u)]
        Qa, ca, Aa = arrs[0], arrs[1], arrs[2]
        count, n = ca.shape
        m = Aa.shape[1]
        if Qa.shape != (count, n, n) or Aa.shape != (count, m, n) or any(
                a.shape != (count, m) for a in arrs[3:5]) or any(a.shape != (count, n) for a in arrs[5:]):
            raise ValueError("Batch.from_dense: inconsistent shapes")
        b._keep = arrs
        b.models = None
        b.ns, b.ms = [n] * count, [m] * count
        h = C.c_void_p()
        st = _LIB.hprqp_hip_batch_create_dense(count, n, m, *[a.ctypes.data_as(_db_p) for a in arrs],
                                               C.byref(b.params.c), C.byref(h))
        if st != STATUS_OK:
            raise ValueError(f"hprqp_hip_batch_create_dense failed: {status_name(st)}")
        b.h = h
        return b

    @classmethod
    def from_dense_device(cls, count: int, n: int, m: int, ptrs: Sequence[int], stream: int = 0,
                          params: Parameters | None = None) -> "Batch":
        """hprqp_hip_batch_create_dense_device: as from_dense, with Q, c, A,
        AL, AU, l, u given as device addresses (float64, contiguous, shapes
        as in from_dense) on params.device_number, read on `stream` (a
        hipStream_t handle as an integer; 0 = the null stream)."""
        b = cls(None, params)
        if len(ptrs) != 7:
            raise ValueError("Batch.from_dense_device: seven device addresses expected")
        b.models = None
        b.ns, b.ms = [n] * count, [m] * count
        h = C.c_void_p()
        st = _LIB.hprqp_hip_batch_create_dense_device(count, n, m, *[C.c_void_p(p) for p in ptrs],
                                                      C.byref(b.params.c), C.c_void_p(stream), C.byref(h))
        if st != STATUS_OK:
            raise ValueError(f"hprqp_hip_batch_create_dense_device failed: {status_name(st)}")
        b.h = h
        return b

    def __len__(self) -> int:
        return len(self.ns)

    def __del__(self) -> None:
        if getattr(self, "h", None):
            self._lib.hprqp_hip_batch_destroy(self.h)
            self.h = None

    def solve(self, bodies_per_launch: int = 100) -> int:
        rem = C.c_size_t(0)
        st = _LIB.hprqp_hip_batch_solve(self.h, bodies_per_launch, C.byref(rem))
        if st != STATUS_OK:
            raise RuntimeError(f"hprqp_hip_batch_solve failed: {status_name(st)}")
        return int(rem.value)

    @property
    def seconds(self) -> float:
        return _LIB.hprqp_hip_batch_seconds(self.h)

    @property
    def wave_row(self) -> int:
        """Row length from which the CSR engine uses a wavefront per row."""
        return _LIB.hprqp_hip_batch_wave_row(self.h)

    def result(self, k: int) -> Result:
        r = CResults()
        _LIB.results_init(C.byref(r))
        st = _LIB.hprqp_hip_batch_result(self.h, k, C.byref(r))
        try:
            if st != STATUS_OK:
                raise RuntimeError(f"hprqp_hip_batch_result failed: {status_name(st)}")
            return _result_of(r, r.status, self.ns[k], self.ms[k])
        finally:
            _LIB.results_free(C.byref(r))

    def solution(self) -> dict:
        """hprqp_hip_batch_solution: x, z (sum n), y (sum m) concatenated in
        original units; status names, iterations, objectives, KKT residuals."""
        sn, sm, cnt = sum(self.ns), sum(self.ms), len(self.ns)
        x, z, y = np.zeros(sn), np.zeros(sn), np.zeros(max(sm, 1))
        status = np.zeros(cnt, np.int32)
        iters = np.zeros(cnt, np.uint64)
        obj, kkt = np.zeros(cnt), np.zeros(cnt)
        st = _LIB.hprqp_hip_batch_solution(self.h, x.ctypes.data_as(_db_p), y.ctypes.data_as(_db_p),
                                           z.ctypes.data_as(_db_p), status.ctypes.data_as(C.POINTER(C.c_int)),
                                           iters.ctypes.data_as(C.POINTER(C.c_size_t)), obj.ctypes.data_as(_db_p),
                                           kkt.ctypes.data_as(_db_p))
        if st != STATUS_OK:
            raise RuntimeError(f"hprqp_hip_batch_solution failed: {status_name(st)}")
        return dict(x=x, y=y[:sm], z=z, status=[status_name(int(v)) for v in status], iter=iters.astype(np.int64),
                    objective=obj, kkt=kkt)

    def status(self) -> dict:
        """Per instance status names, iterations, objectives and KKT
        residuals (hprqp_hip_batch_solution without the vectors)."""
        cnt = len(self.ns)
        status = np.zeros(cnt, np.int32)
        iters = np.zeros(cnt, np.uint64)
        obj, kkt = np.zeros(cnt), np.zeros(cnt)
        st = _LIB.hprqp_hip_batch_solution(self.h, None, None, None, status.ctypes.data_as(C.POINTER(C.c_int)),
                                           iters.ctypes.data_as(C.POINTER(C.c_size_t)), obj.ctypes.data_as(_db_p),
                                           kkt.ctypes.data_as(_db_p))
        if st != STATUS_OK:
            raise RuntimeError(f"hprqp_hip_batch_solution failed: {status_name(st)}")
        return dict(status=[status_name(int(v)) for v in status], iter=iters.astype(np.int64), objective=obj, kkt=kkt)

    def solution_device(self, x: int, y: int, z: int, stream: int = 0) -> None:
        """hprqp_hip_batch_solution_device: x, z (sum n) and y (sum m) in
        original units written to device addresses (0 skips one) by a kernel
        on `stream`, without waiting for it."""
        st = _LIB.hprqp_hip_batch_solution_device(self.h, C.c_void_p(x or None), C.c_void_p(y or None),
                                                  C.c_void_p(z or None), C.c_void_p(stream))
        if st != STATUS_OK:
            raise RuntimeError(f"hprqp_hip_batch_solution_device failed: {status_name(st)}")


def solve_batch(models: Sequence[Model], params: Parameters | None = None, bodies_per_launch: int = 100):
    b = Batch(models, params)
    b.solve(bodies_per_launch)
    return [b.result(k) for k in range(len(b))], b.seconds


_LIB.hprqp_hip_batch_adjoint.argtypes = [C.c_void_p, _db_p, C.c_double, C.c_int, _db_p, _db_p, _db_p,
                                         C.POINTER(C.c_int)]


def batch_adjoint_arrays(batch: Batch, dl_dx: np.ndarray, tol: float = 1e-12, max_steps: int = 20000):
    """Concatenated (u, v_rows, v_bounds, steps) of the active-set KKT adjoint
    K [u; v] = [dl_dx; 0] of every instance (see include/hprqp_hip.h);
    dl_dx is the concatenation of the instances' upstream gradients."""
    sn, sm = sum(batch.ns), sum(batch.ms)
    g = np.ascontiguousarray(dl_dx, dtype=np.float64).reshape(-1)
    assert g.size == sn
    u = np.zeros(sn)
    vb = np.zeros(sn)
    vr = np.zeros(max(sm, 1))
    steps = np.zeros(len(batch.ns), np.int32)
    st = _LIB.hprqp_hip_batch_adjoint(batch.h, g.ctypes.data_as(_db_p), tol, max_steps, u.ctypes.data_as(_db_p),
                                      vr.ctypes.data_as(_db_p), vb.ctypes.data_as(_db_p),
                                      steps.ctypes.data_as(C.POINTER(C.c_int)))
    if st != STATUS_OK:
        raise RuntimeError(f"hprqp_hip_batch_adjoint failed: {status_name(st)}")
    return u, vr[:sm], vb, steps


_LIB.hprqp_hip_batch_create_dense_device.argtypes = ([C.c_size_t] * 3 + [C.c_void_p] * 7
                                                     + [C.POINTER(CParams), C.c_void_p, C.POINTER(C.c_void_p)])
_LIB.hprqp_hip_batch_solution_device.argtypes = [C.c_void_p] * 5
_LIB.hprqp_hip_batch_adjoint_device.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.POINTER(C.c_int), C.c_void_p]


def batch_adjoint_device(batch: Batch, dl_dx: int, u: int, v_rows: int, v_bounds: int, stream: int = 0,
                         tol: float = 1e-12, max_steps: int = 20000, want_steps: bool = False):
    """hprqp_hip_batch_adjoint_device: the adjoint of batch_adjoint_arrays
    with device addresses (dl_dx, u, v_bounds: sum n; v_rows: sum m),
    launched on `stream`; returns the MINRES step counts (waiting for them)
    when want_steps, else None without waiting."""
    steps = np.zeros(len(batch.ns), np.int32) if want_steps else None
    st = _LIB.hprqp_hip_batch_adjoint_device(batch.h, C.c_void_p(dl_dx), tol, max_steps, C.c_void_p(u),
                                             C.c_void_p(v_rows or None), C.c_void_p(v_bounds),
                                             steps.ctypes.data_as(C.POINTER(C.c_int)) if want_steps else None,
                                             C.c_void_p(stream))
    if st != STATUS_OK:
        raise RuntimeError(f"hprqp_hip_batch_adjoint_device failed: {status_name(st)}")
    return steps


_LIB.hprqp_hip_batch_adjoint_residuals.argtypes = [C.c_void_p, _db_p, C.c_void_p]


def batch_adjoint_residuals(batch: Batch, stream: int = 0) -> np.ndarray:
    """Per-instance final relative residual of the batch's last adjoint
    (hprqp_hip_batch_adjoint_residuals), read on `stream` and waited for."""
    out = np.zeros(len(batch.ns))
    st = _LIB.hprqp_hip_batch_adjoint_residuals(batch.h, out.ctypes.data_as(_db_p), C.c_void_p(stream))
    if st != STATUS_OK:
        raise RuntimeError(f"hprqp_hip_batch_adjoint_residuals failed: {status_name(st)}")
    return out


def batch_adjoint(batch: Batch, dl_dx: Sequence[np.ndarray], tol: float = 1e-12, max_steps: int = 20000):
    """Per instance (u, v_rows, v_bounds, steps) of the active-set KKT adjoint
    K [u; v] = [dl_dx; 0] (see include/hprqp_hip.h)."""
    g = np.concatenate([np.asarray(a, np.float64) for a in dl_dx])
    u, vr, vb, steps = batch_adjoint_arrays(batch, g, tol, max_steps)
    out, on, om = [], 0, 0
    for k, (n, m) in enumerate(zip(batch.ns, batch.ms)):
        out.append((u[on:on + n].copy(), vr[om:om + m].copy(), vb[on:on + n].copy(), int(steps[k])))
        on += n
        om += m
    return out
