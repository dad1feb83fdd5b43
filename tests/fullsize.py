"""Helpers of the full-size GPU tests: the committed oracle / reference
fixtures of tests/golden/make_fullsize_golden.py and the comparisons made
against them."""
import hashlib
import json

import numpy as np

import qp_libs as Q

GOLD = Q.ROOT / "tests" / "golden"


def qp_sha256(qp):
    """sha256 of a generated instance's arrays, as the fixture script hashes
    them (the test regenerates the instance and checks it is the same)."""
    h = hashlib.sha256()
    for name in ("q_ptr", "q_ind", "q_val", "a_ptr", "a_ind", "a_val", "c", "AL", "AU", "l", "u"):
        a = getattr(qp, name)
        h.update(np.ascontiguousarray(a, dtype=np.uint64 if a.dtype.kind == "u" else np.float64).tobytes())
    return h.hexdigest()


def fixture(name):
    return json.loads((GOLD / f"{name}.json").read_text())


def rel_err(a, b):
    """max |a - b| / max(1, max |b|): the norm-wise relative error the
    per-iterate parity checks use."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b))))) if b.size else 0.0


def check_sampled(vec, rec, rtol):
    """A full device vector against the fixture's sampled entries and norm
    of the oracle's vector at the same loop body."""
    idx = np.asarray(rec["index"], dtype=np.int64)
    want = np.asarray(rec["value"])
    got = vec[idx]
    scale = max(1.0, float(np.max(np.abs(want))))
    assert np.max(np.abs(got - want)) <= rtol * scale, float(np.max(np.abs(got - want)) / scale)
    assert abs(np.linalg.norm(vec) - rec["norm"]) <= rtol * max(1.0, rec["norm"])
