"""The host polishing hook on the GPU solve (src/shared/solver_backend.c:429,
467-479; src/shared/polish.cpp), with tests/polish_stub_worker.py standing in
for the reference's worker. As the reference's tests/polishing/test_gpu.py:
an accepted candidate is exactly what the solve returns, and failed or slow
workers never change a fixed-length trajectory and are stopped when the
solve ends."""
import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

import qp_libs as Q
from test_gpu_parity import build

pytestmark = pytest.mark.gpu

STUB = Path(__file__).with_name("polish_stub_worker.py")


@pytest.fixture(scope="module")
def H():
    import hprqp_amd
    if not hprqp_amd.gpu_backend_available():
        pytest.fail("no gfx950 device visible: the HIP path must run on the GPU")
    return hprqp_amd


@pytest.fixture(scope="module")
def model(H):
    return build(H, Q.random_qp(seed=0))


def _params(H, tmp, mode, **kw):
    os.environ["HPRQP_TEST_WORKER_MODE"] = mode  # inherited by the worker process
    return H.Parameters(verbose=False, polishing=1, polish_python=sys.executable, polish_worker=str(STUB),
                        polish_directory=str(tmp), polish_keep_files=1, **kw)


def _same(a, b):
    return all(np.array_equal(getattr(a, k), getattr(b, k)) for k in "xyzw")


def test_accepted_candidate_is_the_result(H, model, tmp_path):
    # an unreachable tolerance keeps the GPU iterating until the worker publishes
    r = H.optimize(model, _params(H, tmp_path, "publish", stoptol=1e-300, max_iter=200_000))
    assert r.status == "OPTIMAL" and r.polish_accepted and r.polish_attempts == 1, (r.status, r.polish_attempts)
    assert 700 <= r.iter < 200_000  # first eligible check point: 7 * check_iter
    ready = list(tmp_path.glob("hprqp-polish-*/episode_*/READY"))
    assert len(ready) == 1
    ep = ready[0].parent
    n, m = len(r.x), len(r.y)
    v = np.frombuffer((ep / "candidate.bin").read_bytes()[24:], "<f8")
    assert np.array_equal(r.x, v[:n]) and np.array_equal(r.w, v[n:2 * n]) and np.array_equal(r.z, v[2 * n:3 * n])
    assert np.array_equal(r.y, v[3 * n:3 * n + m])
    assert r.residual == 0.0 and r.objective == 1.0 and r.gap == 0.0  # the certificate's figures
    req = json.loads((ep / "request.json").read_text())
    assert req["dimensions"] == [n, m] and req["tolerance"] == 1e-300 and req["iteration"] == 700
    events = [json.loads(s)["event"] for s in (ep.parent / "events.jsonl").read_text().splitlines()]
    assert events == ["LAUNCHED", "CPU_CERTIFICATE_ACCEPT"]


@pytest.mark.parametrize("mode,iters", [("crash", 3000), ("sleep", 50_000)])  # the sleeper must start before the end
def test_failed_or_slow_worker_leaves_the_trajectory(H, model, tmp_path, mode, iters):
    base = H.optimize(model, H.Parameters(verbose=False, polishing=0, stoptol=1e-300, max_iter=iters))
    r = H.optimize(model, _params(H, tmp_path, mode, stoptol=1e-300, max_iter=iters))
    assert base.status == r.status == "MAX_ITER" and base.iter == r.iter == iters
    assert r.polish_attempts >= 1 and not r.polish_accepted
    assert _same(base, r) and base.objective == r.objective and base.residual == r.residual
    if mode == "crash":
        assert not list(tmp_path.glob("hprqp-polish-*/episode_*/READY"))
    else:
        pids = list(tmp_path.glob("hprqp-polish-*/episode_*/pid"))
        assert len(pids) == 1 and r.polish_attempts == 1  # a pending worker blocks new launches
        with pytest.raises(ProcessLookupError):  # stopped (and reaped) when the solve ended
            os.kill(int(pids[0].read_text()), 0)


def test_attempt_limit_and_no_worker(H, model, tmp_path):
    r = H.optimize(model, _params(H, tmp_path, "crash", stoptol=1e-300, max_iter=20_000, polish_max_attempts=2))
    assert r.polish_attempts == 2 and not r.polish_accepted
    missing = H.Parameters(verbose=False, polishing=1, polish_worker=str(tmp_path / "absent"), stoptol=1e-300,
                           max_iter=1000)
    r = H.optimize(model, missing)
    assert r.status == "MAX_ITER" and r.polish_attempts == 0 and not r.polish_accepted
