"""A stand-in for the reference's CPU polishing worker, for tests only.

It follows the worker side of the protocol the library's owner speaks
(src/shared/polish.cpp:104-131, src/shared/polish/cpu_publication.hpp): read
the request and the snapshot, write candidate.bin, certificate.bin and then
READY. The "polished" candidate is the snapshot itself and the certificate
claims zero residuals; no polishing happens. HPRQP_TEST_WORKER_MODE selects
the behaviour: publish (default), crash (exit status 3) or sleep (write the
process id next to the request, then sleep)."""
import os
import sys
import time
from pathlib import Path


def main() -> int:
    request = Path(sys.argv[1])
    episode = request.parent
    mode = os.environ.get("HPRQP_TEST_WORKER_MODE", "publish")
    if mode == "crash":
        return 3
    if mode == "sleep":
        (episode / "pid").write_text(str(os.getpid()))
        time.sleep(60)
        return 0
    import hashlib
    import json

    import numpy as np

    req = json.loads(request.read_text())
    n, m = req["dimensions"]
    snap = (episode / "snapshot.bin").read_bytes()
    assert snap[:8] == b"HPRSOLS1"
    assert np.frombuffer(snap[8:40], "<u8").tolist() == [n, m, n, n]
    v = np.frombuffer(snap[40:], "<f8")
    x, y, z, w = v[:n], v[n:n + m], v[n + m:2 * n + m], v[2 * n + m:3 * n + m]
    candidate = (b"HPRPOL01" + np.array([n, m], "<u8").tobytes()
                 + np.concatenate([x, w, z, y, np.zeros(m)]).astype("<f8").tobytes())
    (episode / "candidate.bin").write_bytes(candidate)
    tol = float(req["tolerance"])
    certificate = (b"HPRCPU01" + bytes.fromhex(req["token"]) + np.array([n, m], "<u8").tobytes()
                   + hashlib.sha256(candidate).digest()
                   + np.array([tol, 0.0, 0.0, 0.0, 1.0, 1.0], "<f8").tobytes())
    (episode / "certificate.bin").write_bytes(certificate)
    (episode / "READY").write_text(f"HPRCPU01\n{req['token']}\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
