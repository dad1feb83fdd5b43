"""Host polishing hook, CPU side (no GPU): the owner's launch rule, worker
process handling, publication checks and worker lookup (tests/polish_driver.cpp,
compiled here against hpr-qp-c_amd/csrc/polish_owner.cpp), the byte layout
of the files handed to the worker, and the defaults of the reference
(src/shared/solver.c:152-156)."""
import hashlib
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "hpr-qp-c_amd" / "csrc"


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("polish") / "polish_driver"
    subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{CSRC}",
                    str(ROOT / "tests" / "polish_driver.cpp"), str(CSRC / "polish_owner.cpp"), "-o", str(out)],
                   check=True)
    return out


def test_owner_rules_processes_and_certificates(driver, tmp_path):
    payload = tmp_path / "payload.bin"
    payload.write_bytes(np.random.default_rng(5).bytes(200_003))  # several 64 KiB reads, ragged end
    run = tmp_path / "run"
    run.mkdir()
    p = subprocess.run([str(driver), str(run), str(payload)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "failures 0" in p.stdout
    digest = [line.split()[1] for line in p.stdout.splitlines() if line.startswith("sha256 ")][0]
    assert digest == hashlib.sha256(payload.read_bytes()).hexdigest()


def _take(buf, off, fmt, count):
    size = np.dtype("<" + fmt).itemsize * count
    return np.frombuffer(buf[off:off + size], dtype="<" + fmt), off + size


def test_worker_files_layout(driver, tmp_path):
    """model.bin and snapshot.bin as the worker reads them (polish.cpp:33-44,
    113-116): 64-bit sizes, CSR of Q then A, then c, l, u, AL, AU, the cone
    right-hand sides (zeros here), the constant and the cone count; the
    snapshot holds (n, m, n, n) then x, y, z, w."""
    run = tmp_path / "run"
    run.mkdir()
    subprocess.run([str(driver), str(run)], check=True, capture_output=True, timeout=120)
    b = (run / "model.bin").read_bytes()
    assert b[:8] == b"HPRNAT01"
    off = 8
    mats = []
    for _ in range(2):
        dims, off = _take(b, off, "u8", 3)
        nrow, ncol, nnz = (int(v) for v in dims)
        rp, off = _take(b, off, "u8", nrow + 1)
        ci, off = _take(b, off, "u8", nnz)
        val, off = _take(b, off, "f8", nnz)
        mats.append((nrow, ncol, rp.tolist(), ci.tolist(), val.tolist()))
    assert mats[0] == (2, 2, [0, 2, 3], [0, 1, 1], [2.0, 0.5, 3.0])
    assert mats[1] == (1, 2, [0, 2], [0, 1], [1.0, -1.0])
    vecs = {}
    for name, k in (("c", 2), ("l", 2), ("u", 2), ("AL", 1), ("AU", 1), ("rhs", 1), ("const", 1)):
        vecs[name], off = _take(b, off, "f8", k)
    cones, off = _take(b, off, "u8", 1)
    assert off == len(b) and int(cones[0]) == 0
    assert vecs["c"].tolist() == [1.0, -2.0] and vecs["l"][0] == -1.0 and np.isneginf(vecs["l"][1])
    assert vecs["u"][0] == 4.0 and np.isposinf(vecs["u"][1])
    assert vecs["AL"].tolist() == [0.5] and vecs["AU"].tolist() == [2.5] and vecs["rhs"].tolist() == [0.0]
    assert vecs["const"].tolist() == [7.25]
    assert (os.stat(run / "model.bin").st_mode & 0o777) == 0o400
    s = (run / "snapshot.bin").read_bytes()
    assert s[:8] == b"HPRSOLS1"
    dims, off = _take(s, 8, "u8", 4)
    assert dims.tolist() == [2, 1, 2, 2]
    rest = np.frombuffer(s[off:], dtype="<f8")
    assert rest.tolist() == [1.5, -0.5, 0.25, 0.0, 1.0, 3.0, 4.0]  # x, y, z, w


def test_polishing_defaults_match_reference():
    import hprqp_amd as H

    p = H.Parameters().c
    assert p.polishing == 1 and p.polish_max_attempts == 12
    assert p.polish_episode_seconds == 20.0 and p.polish_total_seconds == 0.0
    assert p.polish_memory_bytes == 8 * 1024 ** 3
    assert p.polish_worker is None and p.polish_python is None and p.polish_directory is None
    q = H.Parameters(polish_worker="/opt/w", polish_directory="/tmp").c
    assert q.polish_worker == b"/opt/w" and q.polish_directory == b"/tmp"
