"""Quadratic constraints lifted to row cones (hpr-qp-c_amd/csrc/qcqp.cpp and
the QCMATRIX section of build_from_mps), checked against what the lift must
satisfy and against the reference's own cases:

* the factor lift (src/shared/qcqp.c:16-71; the reference's
  tests/test_qcqp.c): for every x, head^2 - ||tail||^2 = -4 h(x) and the
  cone holds exactly when h(x) <= 0 -- a closed form that a dropped sqrt(2),
  a wrong head sign or a swapped head right-hand side breaks;
* the MPS cases of tests/test_mps_qcqp.cpp:28-117 (cone layout, the factor
  of two between QCMATRIX's x'Hx and the lift's 0.5 x'Dx, the sign of G
  rows, duplicate cards, zero blocks, and every rejected file);
* generated QCMATRIX files, the reference's MPS fixtures and its factor-lift
  data built by this library and by the reference (oracle/_ref) give
  bit-identical models."""
import math

import numpy as np
import pytest

import qp_libs as Q
from test_mps import model_arrays

inf = math.inf


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    return hprqp_amd


def csr(nrow, ncol, rp, ci, v):
    return (nrow, ncol, np.array(rp, np.uint64), np.array(ci, np.uint64), np.array(v, np.float64))


# tests/test_qcqp.c:10-12: two constraints, factors of two and one row
C_FAC = csr(2, 3, [0, 2, 3], [0, 2, 1], [2, -1, .5])
R_FAC = csr(3, 3, [0, 2, 4, 5], [0, 1, 1, 2, 0], [1, 2, -3, 4, 2])
OFF = [0, 2, 3]
B = [-3.0, 2.0]


def dense(M):
    nrow, ncol, rp, ci, v = M
    D = np.zeros((nrow, ncol))
    for i in range(nrow):
        for k in range(int(rp[i]), int(rp[i + 1])):
            D[i, int(ci[k])] += v[k]
    return D


def factor_model(H, R=R_FAC):
    return H.build_from_qabc_qc(np.zeros((3, 3)), np.zeros(3), np.zeros((0, 3)), [], [], [-inf] * 3, [inf] * 3, 7.0,
                                C_FAC, B, R=R, offsets=OFF)


def test_factor_lift_identity(H):
    qp = model_arrays(factor_model(H))
    assert qp.n == 3 and qp.m == 7 and len(qp.a_val) == 2 * 3 + 5
    assert list(qp.row_offsets) == [0, 4, 7]
    assert qp.obj_constant == 7.0
    A, Cd, Rd = qp.dense_A(), dense(C_FAC), dense(R_FAC)
    for t in range(-20, 21):
        x = np.array([t * .07, t * .03 + .1, .2 - t * .02])
        for i in range(2):
            h = Cd[i] @ x + B[i] + .5 * sum((Rd[j] @ x) ** 2 for j in range(OFF[i], OFF[i + 1]))
            first, end = int(qp.row_offsets[i]), int(qp.row_offsets[i + 1])
            s = A[first:end] @ x - qp.soc_rhs[first:end]
            head, sq = s[0], float(np.sum(s[1:] ** 2))
            assert abs(head * head - sq + 4 * h) < 1e-11 * (1 + abs(h))
            assert (h <= 0) == (head >= math.sqrt(sq))
    assert np.all(np.isinf(qp.AL)) and np.all(qp.AL < 0) and np.all(np.isinf(qp.AU))


def test_factor_lift_rejects_nan(H):
    bad = list(R_FAC)
    bad[4] = bad[4].copy()
    bad[4][0] = np.nan
    with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
        factor_model(H, R=tuple(bad))


def test_zero_rank_factor_and_no_constraints(H):
    """A constraint with no factor rows is a 2-row cone (the linear
    inequality C x + k <= 0); no constraints at all is build_from_QAbc."""
    R0 = csr(1, 3, [0, 1], [0], [1.0])
    m = H.build_from_qabc_qc(np.zeros((3, 3)), np.ones(3), np.zeros((0, 3)), [], [], [-1] * 3, [1] * 3, 0.0,
                             C_FAC, B, R=R0, offsets=[0, 0, 1])
    qp = model_arrays(m)
    assert list(qp.row_offsets) == [0, 2, 5]
    none = csr(0, 3, [0], [], [])
    m = H.build_from_qabc_qc(np.eye(3), np.ones(3), [[1.0, 1.0, 0.0]], [-1.0], [1.0], [-1] * 3, [1] * 3, 0.0,
                             none, [], R=none, offsets=[0])
    qp = model_arrays(m)
    assert qp.num_row_cones == 0 and qp.m == 1


def test_diagonal_lift_equals_factor_lift(H):
    """D_ij > 0 becomes the factor row sqrt(D_ij) e_j: the diagonal form and
    the explicit factors give the same model bit for bit."""
    D = csr(2, 3, [0, 2, 3], [0, 2, 1], [4.0, 0.0, 2.5])
    R = csr(2, 3, [0, 1, 2], [0, 1], [2.0, math.sqrt(2.5)])
    a = model_arrays(H.build_from_qabc_qc(np.zeros((3, 3)), np.zeros(3), np.zeros((0, 3)), [], [], [-inf] * 3,
                                          [inf] * 3, 0.0, C_FAC, B, D=D))
    b = model_arrays(H.build_from_qabc_qc(np.zeros((3, 3)), np.zeros(3), np.zeros((0, 3)), [], [], [-inf] * 3,
                                          [inf] * 3, 0.0, C_FAC, B, R=R, offsets=[0, 1, 2]))
    for f in ("a_ptr", "a_ind", "a_val", "soc_rhs", "row_offsets"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    neg = csr(2, 3, [0, 1, 1], [0], [-1.0])
    with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
        H.build_from_qabc_qc(np.zeros((3, 3)), np.zeros(3), np.zeros((0, 3)), [], [], [-inf] * 3, [inf] * 3, 0.0,
                             C_FAC, B, D=neg)


# ---------------------------------------------------------------- MPS cases

def fixture(sense, linear, rhs, qc, extra=""):
    """tests/test_mps_qcqp.cpp:28-34: min -x + 3 s.t. x in [0, 4], y = 0 and
    one quadratic row."""
    return ("NAME QC\nROWS\n N OBJ\n " + sense + " QC\nCOLUMNS\n X OBJ -1 QC " + linear + "\n Y OBJ 0\nRHS\n RHS1 OBJ -3 QC "
            + rhs + "\nBOUNDS\n UP BND X 4\n FX BND Y 0\n" + extra + "QCMATRIX QC\n" + qc + "ENDATA\n")


def load(H, tmp_path, text, name="qc.mps"):
    p = tmp_path / name
    p.write_text(text)
    return H.read_from_mps(p)


def entry(qp, r, c):
    return qp.dense_A()[r, c]


@pytest.mark.parametrize("text", [fixture("L", "1", "4", " X X .5\n"), fixture("G", "-1", "-4", " X X -.5\n"),
                                  fixture("L", "1", "4", " X X .25\n X X 2.5D-1\n")], ids=["L", "G", "repeated"])
def test_mps_quadratic_row_cone(H, tmp_path, text):
    """x'Hx with H = .5 is 0.5*x'Dx with D = 1; the cone is
    (5 - x, 3 - x, sqrt(2) x), equivalent to .5x^2 + x <= 4."""
    qp = model_arrays(load(H, tmp_path, text))
    assert qp.num_row_cones == 1 and qp.m == 3
    assert qp.n == 2 and len(qp.q_val) == 0
    assert qp.c[0] == -1 and qp.obj_constant == 3
    assert qp.l[0] == 0 and qp.u[0] == 4
    assert entry(qp, 0, 0) == -1 and entry(qp, 1, 0) == -1
    assert abs(entry(qp, 2, 0) - math.sqrt(2.0)) < 1e-14
    assert qp.soc_rhs[0] == -5 and qp.soc_rhs[1] == -3
    for x in (-1., 0., 1., 2., 3., 4.):
        t, v, tail = 5 - x, 3 - x, math.sqrt(2.0) * x
        assert (t >= math.hypot(v, tail) - 1e-13) == (.5 * x * x + x <= 4)


def test_mps_mixed_rows_keep_order(H, tmp_path):
    """Duplicate linear cards are summed and the linear rows stay ahead of
    the cone (tests/test_mps_qcqp.cpp:68-76)."""
    qp = model_arrays(load(H, tmp_path, "NAME DUP\nROWS\n N OBJ\n E LIN\n L QC\nCOLUMNS\n X OBJ -1 LIN 1\n"
                                        " X LIN 2 QC 1\n Y QC 2\nRHS\n R LIN 3 QC 4\nQCMATRIX QC\n Y Y .5\nENDATA\n"))
    assert int(qp.row_offsets[0]) == 1
    assert entry(qp, 0, 0) == 3 and entry(qp, 1, 0) == -1 and entry(qp, 1, 1) == -2
    assert qp.AL[0] == 3 and qp.AU[0] == 3


def test_mps_zero_quadratic_block_stays_linear(H, tmp_path):
    qp = model_arrays(load(H, tmp_path, fixture("E", "1", "2", " X X 0\n")))
    assert qp.num_row_cones == 0 and qp.m == 1 and qp.AU[0] == 2


def _rejected_texts():
    base = fixture("L", "1", "4", " X X .5\n")
    out = {f"qc{i}": fixture("L", "1", "4", qc) for i, qc in enumerate(
        [" X Y 1\n", " Z Z 1\n", " X X nan\n", " X X inf\n", " X X bad\n", " X X 1e999\n", " X X\n", " X X -.5\n",
         " X X .5\nQCMATRIX QC\n X X .5\n"])}
    out.update({"equality": fixture("E", "1", "4", " X X .5\n"), "g_convex": fixture("G", "1", "4", " X X .5\n"),
                "ranged": fixture("L", "1", "4", " X X .5\n", "RANGES\n R QC 2\n"),
                "linear_oops": fixture("L", "oops", "4", " X X .5\n"),
                "rhs_oops": fixture("L", "1", "oops", " X X .5\n"),
                "indicators": fixture("L", "1", "4", " X X .5\n", "INDICATORS\n IF QC X 1\n"),
                "unknown_row": base.replace("QCMATRIX QC", "QCMATRIX ZZ"),
                "no_endata": base[:base.find("ENDATA")],
                "binary": base.replace("UP BND X 4", "BV BND X")})
    return out


REJECTED = _rejected_texts()


@pytest.mark.parametrize("name", sorted(REJECTED))
def test_mps_quadratic_rejected(H, tmp_path, name):
    with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
        load(H, tmp_path, REJECTED[name])
    if Q.ref_available():
        p = tmp_path / "ref.mps"
        p.write_text(REJECTED[name])
        assert Q.binding_solve(Q.ref_lib(), None, mps_path=p, use_gpu=0).get("build_failed")


def qc_mps(path, seed, n=30):
    """A generated model with quadratic rows: some L rows (H >= 0), some G
    rows (H <= 0, listed with a repeated card), one block of zeros."""
    qp = Q.random_qp(seed=seed, n=n, m_eq=6, m_in=10, a_row_nnz=5, b_rows=8, b_row_nnz=3)
    rng = np.random.default_rng(seed)
    qc = {}
    for k, i in enumerate(range(6, 16)):
        if k % 3 == 0:
            qp.AL[i] = -inf
            cols = rng.choice(n, 3, replace=False)
            qc[i] = [(int(j), float(rng.uniform(.1, 2))) for j in cols]
        elif k % 3 == 1:
            qp.AU[i] = inf
            j = int(rng.integers(n))
            qc[i] = [(j, -0.25), (int(rng.integers(n)), -1.5), (j, -0.5)]
    qp.AL[6 + 2] = -inf
    qc[6 + 2] = [(0, 0.0)]
    Q.write_mps(qp, path, seed=seed, qc=qc)
    return qc


def test_generated_quadratic_rows_model(H, tmp_path):
    """The lift of a generated file, checked row by row against the cards."""
    path = tmp_path / "gqc.mps"
    qc = qc_mps(path, seed=11)
    qp = model_arrays(H.read_from_mps(path))
    lifted = [i for i, cards in qc.items() if any(h != 0.0 for _, h in cards)]
    assert qp.num_row_cones == len(lifted)
    for k, i in enumerate(sorted(lifted)):
        b, e = int(qp.row_offsets[k]), int(qp.row_offsets[k + 1])
        agg = {}
        for j, h in qc[i]:
            agg[j] = agg.get(j, 0.0) + h
        nz = sorted(j for j, h in agg.items() if h != 0.0)
        assert e - b == 2 + len(nz)
        A = qp.dense_A()
        for t, j in enumerate(nz):
            row = A[b + 2 + t]
            assert np.count_nonzero(row) == 1
            assert row[j] == pytest.approx(math.sqrt(2.0) * math.sqrt(2.0 * abs(agg[j])), rel=1e-15)
        np.testing.assert_array_equal(A[b], A[b + 1])
        assert qp.soc_rhs[b + 1] - qp.soc_rhs[b] == 2.0


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_quadratic_rows_equal_reference_loader(H, tmp_path, seed):
    """Bitwise the reference loader's model. (The reference's CPU solver
    returns NOT_IMPLEMENTED on cone models, src/shared/solver.c:250-253, so
    the lifted models are compared as models; the GPU tests solve them
    against the oracle.)"""
    path = tmp_path / f"qc{seed}.mps"
    qc_mps(path, seed=seed)
    mine = model_arrays(H.read_from_mps(path))
    assert mine.num_row_cones > 0
    Q.assert_same_model(mine, Q.ref_model_from_mps(path))


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("text", [fixture("L", "1", "4", " X X .5\n"), fixture("G", "-1", "-4", " X X -.5\n"),
                                  fixture("E", "1", "2", " X X 0\n")], ids=["L", "G", "zero"])
def test_fixture_models_equal_reference_loader(H, tmp_path, text):
    p = tmp_path / "f.mps"
    p.write_text(text)
    Q.assert_same_model(model_arrays(H.read_from_mps(p)), Q.ref_model_from_mps(p))


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
def test_factor_lift_equals_reference(H):
    """build_from_QAbc_qc_factors of both libraries on the reference's own
    test data (tests/test_qcqp.c:10-13)."""
    mine = model_arrays(factor_model(H))
    L = Q.ref_lib()
    import ctypes as C
    keep = []

    def mat(M):
        nrow, ncol, rp, ci, v = M
        rp, ci, v = (np.ascontiguousarray(a if len(a) else np.zeros(1, a.dtype)) for a in (rp, ci, v))
        keep.extend([rp, ci, v])
        return Q._CSR(nrow, ncol, int(rp[-1]), rp.ctypes.data_as(Q.c_size_p), ci.ctypes.data_as(Q.c_size_p),
                      v.ctypes.data_as(Q.c_double_p), 0)

    def dv(a):
        a = np.ascontiguousarray(a, np.float64)
        keep.append(a)
        return a.ctypes.data_as(Q.c_double_p)
    Qz, Az = mat(csr(3, 3, [0, 0, 0, 0], [], [])), mat(csr(0, 3, [0], [], []))
    off = np.array(OFF, np.uint64)
    m = Q._CModel()
    st = L.build_from_QAbc_qc_factors(C.byref(Qz), C.byref(Az), dv(np.zeros(3)), None, None, dv([-inf] * 3),
                                      dv([inf] * 3), C.c_double(7.0), C.byref(mat(C_FAC)), dv(B), C.byref(mat(R_FAC)),
                                      off.ctypes.data_as(Q.c_size_p), C.byref(m))
    assert st == 0
    try:
        Q.assert_same_model(mine, Q.qp_of_cmodel(m))
    finally:
        L.model_free(C.byref(m))


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
def test_random_factor_lifts_agree_with_reference(H):
    """build_from_QAbc_qc_factors of this library and of the reference on
    random small inputs (factor blocks of any rank, a share of them invalid:
    NaN factors, bad offsets, non-finite constants): the same status, and
    bitwise-equal models when both accept."""
    import ctypes as C
    L = Q.ref_lib()
    rng = np.random.default_rng(5)

    def rand_csr(nrow, ncol, per_row):
        ptr, ind, val = [0], [], []
        for _ in range(nrow):
            cols = np.sort(rng.choice(ncol, size=int(rng.integers(0, min(per_row, ncol) + 1)), replace=False))
            ind += [int(c) for c in cols]
            val += [float(v) for v in rng.standard_normal(len(cols))]
            ptr.append(len(ind))
        return csr(nrow, ncol, ptr, ind, val)

    ok = 0
    for it in range(200):
        n, nq = int(rng.integers(1, 6)), int(rng.integers(0, 4))
        offs = np.sort(rng.integers(0, 4, size=nq + 1)) if nq else np.zeros(1, np.int64)
        offs = np.cumsum(np.concatenate([[0], np.diff(offs)])).astype(np.uint64) if nq else np.zeros(1, np.uint64)
        R = rand_csr(int(offs[-1]), n, 3)
        Cm = rand_csr(nq, n, 3)
        const = rng.standard_normal(nq)
        damage = int(rng.integers(6))
        if damage == 1 and len(R[4]):
            R[4][0] = np.nan
        elif damage == 2 and nq:
            offs = offs.copy()
            offs[-1] += 1  # does not end at R's row count
        elif damage == 3 and nq:
            const[0] = np.inf
        A = rand_csr(int(rng.integers(0, 3)), n, 2)
        AL = -np.abs(rng.standard_normal(A[0])) - 0.1
        AU = np.abs(rng.standard_normal(A[0])) + 0.1
        lb, ub = -np.ones(n), np.ones(n)
        try:
            mine = model_arrays(H.build_from_qabc_qc(np.eye(n), np.ones(n), A, AL, AU, lb, ub, 0.5, Cm, const, R=R,
                                                     offsets=offs))
        except ValueError:
            mine = None
        keep = []

        def mat(M):
            nrow, ncol, rp, ci, v = M
            rp, ci, v = (np.ascontiguousarray(a if len(a) else np.zeros(1, a.dtype)) for a in (rp, ci, v))
            keep.extend([rp, ci, v])
            return Q._CSR(nrow, ncol, int(rp[-1]) if nrow else 0, rp.ctypes.data_as(Q.c_size_p),
                          ci.ctypes.data_as(Q.c_size_p), v.ctypes.data_as(Q.c_double_p), 0)

        def dv(a):
            a = np.ascontiguousarray(a if len(a) else np.zeros(1), np.float64)
            keep.append(a)
            return a.ctypes.data_as(Q.c_double_p)
        eye = csr(n, n, list(range(n + 1)), list(range(n)), [1.0] * n)
        off = np.ascontiguousarray(offs, np.uint64)
        m = Q._CModel()
        st = L.build_from_QAbc_qc_factors(C.byref(mat(eye)), C.byref(mat(A)), dv(np.ones(n)), dv(AL), dv(AU), dv(lb),
                                          dv(ub), C.c_double(0.5), C.byref(mat(Cm)), dv(const), C.byref(mat(R)),
                                          off.ctypes.data_as(Q.c_size_p), C.byref(m))
        try:
            ref = Q.qp_of_cmodel(m) if st == 0 else None
        finally:
            L.model_free(C.byref(m))
        assert (mine is None) == (ref is None), (it, st, damage)
        if mine is not None:
            Q.assert_same_model(mine, ref)
            ok += 1
    assert 40 < ok < 200
