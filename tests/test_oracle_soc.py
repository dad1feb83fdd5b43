"""Cone (SOC) part of the oracle. The reference has no CPU path for cones
(src/shared/solver.c:250-253), so the restatement of the GPU semantics
(src/gpu/soc_*.cuh) is pinned by the reference's own analytic SOC cases and
certificate (tests/gpu/test_gpu_soc.cu:152-273): Euclidean projection of
(0,3,4) onto the cone -> (2.5,1.5,2); min t with (t-2,3,4) in the cone -> 7;
the constant affine cone (feasible -> x=1, infeasible -> MAX_ITER with
residual >= 0.49); and the mixed 8x7 model whose optimum is constructed
(x = (0.25,-0.5,5,3,4,2,0,0)).
"""
import numpy as np
import pytest

import qp_libs as Q


@pytest.mark.parametrize("scaling", [0, 1])
@pytest.mark.parametrize("name", ["proj_cone", "affine_cone", "constant_cone", "constant_cone_infeasible"])
def test_soc_simple_models(name, scaling):
    qp, want = Q.soc_simple_models()[name]
    kw = dict(use_scaling=scaling, stoptol=1e-9 if name != "constant_cone" else 1e-8, print_frequency=0,
              max_iter=100000 if name != "constant_cone_infeasible" and name != "constant_cone" else 1000)
    o = Q.OracleSession(qp, **kw)
    assert o.status == 0
    r, x, y, z, w = o.solve()
    if want is None:
        assert Q.STATUS[r.status] == "MAX_ITER" and r.residuals >= 0.49
        return
    assert Q.STATUS[r.status] == "OPTIMAL", (name, r.status, r.iter, r.residuals)
    assert np.max(np.abs(x - want)) <= 2e-7 if name != "constant_cone" else 1e-7
    cert = Q.soc_certificate(qp, x, y, z)
    assert max(cert.values()) < 2e-6, cert


@pytest.mark.parametrize("q_kind", [0, 1, 2])
@pytest.mark.parametrize("scaling", [0, 1])
def test_soc_mixed_model(q_kind, scaling):
    qp = Q.soc_mixed_model(q_kind)
    o = Q.OracleSession(qp, use_scaling=scaling, stoptol=1e-9, print_frequency=0, max_iter=100000, time_limit=30.0)
    r, x, y, z, w = o.solve()
    assert Q.STATUS[r.status] == "OPTIMAL", (r.status, r.iter, r.residuals)
    cert = Q.soc_certificate(qp, x, y, z)
    assert max(cert.values()) < 2e-6, cert
    if q_kind:
        assert np.max(np.abs(x - [0.25, -0.5, 5, 3, 4, 2, 0, 0])) <= 2e-5
