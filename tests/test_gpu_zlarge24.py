"""FP64 parity at 24^4 (the single-large-lattice configuration) against the
reference program's own output: the first 50 RK3 steps (eps 0.01, E and Q
every 5 steps) of one 24^4, beta 6.0 configuration, and q(x) of the flowed
configuration (MeasureQdensityDistribution, LGF/LGF.cpp:114-124), as
recorded in tests/golden/ref_large24.json by the reference built from its
sources (LGF/LGF.cpp:163-196).

The input is regenerated bit for bit by the CPU oracle's seeded generator
(cold start, 200 heatbath/overrelaxation iterations); that takes minutes of
CPU, so it runs in a background process started when collection finishes
(tests/large24_input.py, conftest.py), and this file sorts after the other
GPU tests so that the generation overlaps them.  Its sha256 must equal the
one recorded with the reference's outputs.
"""
import numpy as np
import pytest

from msc_project_amd import lgf
from tests import large24_input
from tests.common import TOL_FP64, flow_measure, golden, q_err, q_rel_strict, rel_err

pytestmark = pytest.mark.gpu


def test_large24_golden_fp64():
    d = golden("large24")
    shape = tuple(d["shape"])
    U, digest, log = large24_input.wait()
    print(log.strip().splitlines()[-1], flush=True)
    assert digest == d["input_sha256"][0]
    c = d["configs"][0]
    ctx = lgf.Context(0)
    f = lgf.Field(ctx, shape, lgf.FP64)
    f.upload(U)
    del U
    t, E, Q, Qa = flow_measure(f, d["eps"], d["steps"], d["every"])
    strict, nstrict = q_rel_strict(Q, c["Q"], Qa)
    print("24^4: max rel dE %.2e, max dQ/scale %.2e, max |dQ|/|Q| %.2e over %d of %d measurements" %
          (rel_err(E, c["E"]), q_err(Q, c["Q"], Qa), strict, nstrict, len(Q)), flush=True)
    assert np.allclose(t, c["t"], rtol=0, atol=1e-12)
    assert rel_err(E, c["E"]) < TOL_FP64
    assert q_err(Q, c["Q"], Qa) < TOL_FP64
    assert strict < TOL_FP64
    # q(x) of the flowed configuration, in the reference's file scale (pi^2/4)
    o = f.observables(lgf.MEASURE_DEFAULT, want_q=True)
    q = o["qraw"] * (np.pi ** 2 / 4)
    scale = c["qx_abs_sum"]
    assert abs(q.sum() - c["qx_sum"]) < TOL_FP64 * scale
    for k, v in c["qx_samples"].items():
        assert abs(q[int(k)] - v) < TOL_FP64 * max(abs(v), scale / q.size), k
    f.close()
    ctx.close()
