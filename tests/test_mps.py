"""build_from_mps (hpr-qp-c_amd/csrc/mps.cpp) against the reference's MPS
loader (src/shared/mps_reader.cpp, src/shared/mps_qp.cpp): the reference's
own MPS files (tests/data/simple_qp.mps, model.mps, kept as fixtures under
tests/golden/mps) with the models written out by hand, and generated files
exercising every supported section and card, read by both loaders and solved
on the CPU -- the oracle on this library's model, the reference on its own
-- with bit-identical iterates required (identical models)."""
import ctypes as C
import math

import numpy as np
import pytest

import qp_libs as Q

GOLD = Q.ROOT / "tests" / "golden" / "mps"
inf = math.inf


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    return hprqp_amd


def model_arrays(model):
    """The preprocessed model as a qp_libs.QP (cones included)."""
    return Q.qp_of_cmodel(model.c)


def test_simple_qp_fixture_model(H):
    qp = model_arrays(H.read_from_mps(GOLD / "simple_qp.mps"))
    assert (qp.n, qp.m) == (2, 2)
    np.testing.assert_array_equal(qp.c, [1.0, 2.0])
    np.testing.assert_array_equal(qp.dense_A(), [[1.0, 1.0], [1.0, -1.0]])
    np.testing.assert_array_equal(qp.AL, [1.0, -inf])
    np.testing.assert_array_equal(qp.AU, [inf, 2.0])
    np.testing.assert_array_equal(qp.l, [0.0, -5.0])
    np.testing.assert_array_equal(qp.u, [10.0, 5.0])
    np.testing.assert_array_equal(qp.dense_Q(), [[2.0, 1.0], [1.0, 4.0]])  # one triangle, mirrored
    assert qp.obj_constant == 0.0


def test_model_mps_fixture_model(H):
    qp = model_arrays(H.read_from_mps(GOLD / "model.mps"))
    np.testing.assert_array_equal(qp.c, [-3.0, -5.0])
    np.testing.assert_array_equal(qp.dense_A(), [[1.0, 2.0], [3.0, 1.0]])
    np.testing.assert_array_equal(qp.AL, [-inf, -inf])
    np.testing.assert_array_equal(qp.AU, [10.0, 12.0])
    np.testing.assert_array_equal(qp.l, [0.0, 0.0])
    np.testing.assert_array_equal(qp.u, [inf, inf])  # PL
    np.testing.assert_array_equal(qp.dense_Q(), [[2.0, 0.0], [0.0, 2.0]])


def test_generated_file_semantics(H, tmp_path):
    """Ranges, bound kinds and the objective constant as the format defines
    them, checked against values computed here from the written cards."""
    qp = Q.random_qp(seed=4, n=16, m_eq=5, m_in=7, a_row_nnz=4, b_rows=4, b_row_nnz=3)
    path = tmp_path / "gen.mps"
    Q.write_mps(qp, path, seed=1)
    got = model_arrays(H.read_from_mps(path))
    assert got.obj_constant == 1.25  # RHS of -1.25 on the objective row
    names = sorted(Q._BOUND_KINDS)
    want = {"default": (0.0, inf), "box": (-2.0, 3.0), "fixed": (0.5, 0.5), "free": (-inf, inf),
            "minus": (-inf, 4.0), "up_neg": (-inf, -1.5), "up_pos": (0.0, 2.5), "lo_only": (-3.0, inf)}
    for j in range(qp.n):
        assert (got.l[j], got.u[j]) == want[names[j % len(names)]], j
    np.testing.assert_array_equal(got.c, qp.c)  # the second objective row is ignored
    # rows: equalities keep AL = AU unless ranged; ranged L/G rows reproduce the interval
    for i in range(qp.m):
        lo, hi = qp.AL[i], qp.AU[i]
        if lo == hi:
            if i % 7 != 3:
                assert got.AL[i] == got.AU[i] == lo
        elif not (math.isinf(lo) or math.isinf(hi)):
            assert got.AL[i] == pytest.approx(lo, abs=1e-12) and got.AU[i] == pytest.approx(hi, abs=1e-12)


def test_linear_program_file(H, tmp_path):
    """No QUADOBJ section: Q is the empty n x n matrix."""
    path = tmp_path / "lp.mps"
    path.write_text("NAME LP\nROWS\n N obj\n L r\n G s\nCOLUMNS\n x obj 1 r 1\n y obj -1 r 1\n y s 2\nRHS\n rhs r 4 s 1\n"
                    "BOUNDS\n UP b y 3\nENDATA\n")
    qp = model_arrays(H.read_from_mps(path))
    assert (qp.n, qp.m, len(qp.q_val)) == (2, 2, 0)
    np.testing.assert_array_equal(qp.dense_A(), [[1.0, 1.0], [0.0, 2.0]])
    if Q.ref_available():
        Q.assert_same_model(qp, Q.ref_model_from_mps(path))


def test_empty_model_file(H, tmp_path):
    """NAME, ROWS and ENDATA alone: a model without variables or rows, as the
    reference loader builds it."""
    path = tmp_path / "empty.mps"
    path.write_text("NAME E\nROWS\nENDATA\n")
    qp = model_arrays(H.read_from_mps(path))
    assert (qp.n, qp.m) == (0, 0)
    if Q.ref_available():
        Q.assert_same_model(qp, Q.ref_model_from_mps(path))


REF_CASES = [("fixture", "simple_qp.mps", None), ("fixture", "model.mps", None),
             ("gen", "QUADOBJ", 1), ("gen", "QMATRIX", 2), ("gen", "QUADOBJ", 3)]


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
@pytest.mark.parametrize("kind,arg,seed", REF_CASES)
def test_model_equals_reference_loader(H, tmp_path, kind, arg, seed):
    if kind == "fixture":
        path = GOLD / arg
    else:
        qp0 = Q.random_qp(seed=seed, n=40, m_eq=8, m_in=12, a_row_nnz=5, b_rows=10, b_row_nnz=3)
        path = tmp_path / f"gen{seed}.mps"
        Q.write_mps(qp0, path, seed=seed, quad=arg)
    qp = model_arrays(H.read_from_mps(path))
    Q.assert_same_model(qp, Q.ref_model_from_mps(path))
    for k in (3, 40, 300):
        extra = {"max_iter": k}
        r = Q.binding_solve(Q.ref_lib(), None, mps_path=path, use_gpu=0, **extra)
        assert not r.get("build_failed"), r
        o = Q.OracleSession(qp, **extra)
        res, x, y, z, w = o.solve()
        assert Q.STATUS[res.status] == r["status"] and res.iter == r["iter"], (k, r["status"])
        assert res.primal_obj == r["objective"]
        for v, arr in zip("xyzw", (x, y, z, w)):
            np.testing.assert_array_equal(arr, r[v], err_msg=f"{v} at max_iter {k}")


BAD = {
    "maximise": "NAME X\nOBJSENSE\n MAX\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\nRHS\n rhs r 1\nENDATA\n",
    "integer": ("NAME X\nROWS\n N obj\n L r\nCOLUMNS\n m 'MARKER' 'INTORG'\n x obj 1 r 1\n m 'MARKER' 'INTEND'\n"
                "RHS\n rhs r 1\nENDATA\n"),
    "binary_bound": "NAME X\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\nBOUNDS\n BV b x\nENDATA\n",
    "unknown_row": "NAME X\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 q 1\nENDATA\n",
    "no_endata": "NAME X\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\n",
    "bad_number": "NAME X\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1x\nENDATA\n",
    "unknown_bound": "NAME X\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\nBOUNDS\n ZZ b x 1\nENDATA\n",
    "asymmetric_qmatrix": ("NAME X\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1\n y obj 1 r 1\nQMATRIX\n x y 1\n"
                           " y x 2\nENDATA\n"),
    "duplicate_row": "NAME X\nROWS\n N obj\n L r\n G r\nCOLUMNS\n x obj 1 r 1\nENDATA\n",
    "long_line": "NAME X\nROWS\n N obj\n L r\nCOLUMNS\n x obj 1 r 1" + " " * 1100 + "\nENDATA\n",
    "columns_first": "NAME X\nCOLUMNS\n x obj 1\nROWS\n N obj\nENDATA\n",
}


@pytest.mark.parametrize("name", sorted(BAD))
def test_rejected_files(H, tmp_path, name):
    path = tmp_path / f"{name}.mps"
    path.write_text(BAD[name])
    with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
        H.read_from_mps(path)
    if Q.ref_available():
        r = Q.binding_solve(Q.ref_lib(), None, mps_path=path, use_gpu=0)
        assert r.get("build_failed"), "the reference loader rejects it too"


def test_missing_file(H, tmp_path):
    with pytest.raises(ValueError, match="INVALID_ARGUMENT"):
        H.read_from_mps(tmp_path / "none.mps")


MUTANT_TOKENS = ["1e999", "nan", "x", "1.5", "-0", "1d2", "+3", ".5e1", "inf", "1e-400"]
MUTANT_LINES = ["RANGES", "BOUNDS", "RHS", "OBJSENSE", " MAX", " MIN", "OBJSENSE MIN", "ENDATA", "*c", "", "NAME B",
                "QCMATRIX R1", " X1 X1 0.5", "QCMATRIX R6", " X2 X3 1", " X2 X2 -1", "QUADOBJ", " RNG R7 1"]


def mutate(lines, rng):
    """One to three random edits of an MPS file's lines: deleted, repeated,
    inserted or re-indented lines, replaced, dropped or extra fields, tabs."""
    lines = list(lines)
    for _ in range(rng.integers(1, 4)):
        i = int(rng.integers(len(lines)))
        toks, lead = lines[i].split(), (" " if lines[i].startswith((" ", "\t")) else "")
        op = int(rng.integers(9))
        if op == 0 and len(lines) > 3:
            del lines[i]
        elif op == 1:
            lines.insert(i, lines[int(rng.integers(len(lines)))])
        elif op == 2 and toks:
            toks[int(rng.integers(len(toks)))] = str(rng.choice(MUTANT_TOKENS))
            lines[i] = lead + " ".join(toks)
        elif op == 3 and len(toks) > 1:
            del toks[int(rng.integers(len(toks)))]
            lines[i] = lead + " ".join(toks)
        elif op == 4:
            lines[i] = lines[i].strip()
        elif op == 5:
            lines[i] = "   " + lines[i]
        elif op == 6:
            lines.insert(i, str(rng.choice(MUTANT_LINES)))
        elif op == 7 and toks:
            lines[i] = lead + " ".join(toks + [str(rng.choice(["R0", "X1", "1.0", "BND"]))])
        elif op == 8:
            lines[i] = lines[i].replace(" ", "\t", 1)
    return lines


@pytest.mark.skipif(not Q.ref_available(), reason="reference not built here (oracle/_ref)")
def test_mutated_files_agree_with_reference_loader(H, tmp_path):
    """Randomly edited MPS files (generated QUADOBJ, QMATRIX and QCMATRIX
    models): this library and the reference loader accept and reject the
    same files, and build bitwise-equal models from the accepted ones."""
    from test_qcqp import qc_mps
    bases = []
    for seed in range(4):
        qp0 = Q.random_qp(seed=seed, n=12, m_eq=3, m_in=4, a_row_nnz=3, b_rows=3, b_row_nnz=2)
        p = tmp_path / f"base{seed}.mps"
        Q.write_mps(qp0, p, seed=seed, quad=("QUADOBJ", "QMATRIX")[seed % 2])
        bases.append(p.read_text().splitlines())
    for seed in range(2):
        p = tmp_path / f"qc{seed}.mps"
        qc_mps(p, seed=30 + seed, n=10)
        bases.append(p.read_text().splitlines())
    rng = np.random.default_rng(2024)
    accepted = 0
    for it in range(240):
        p = tmp_path / f"m{it}.mps"
        p.write_text("\n".join(mutate(bases[int(rng.integers(len(bases)))], rng)) + "\n")
        try:
            mine = model_arrays(H.read_from_mps(p))
        except ValueError:
            mine = None
        ref = Q.ref_model_from_mps(p)
        assert (mine is None) == (ref is None), (it, "accepted by " + ("the reference" if mine is None else "this library"))
        if mine is not None:
            Q.assert_same_model(mine, ref)
            accepted += 1
    assert accepted > 30  # the edits leave a fair share of files valid
