"""Large-volume q(x) case: one 64^4 lattice at beta 6.2 from the seeded GPU
heatbath, flowed with RK3 (eps 0.01) to t = 8 in the mixed-precision mode
(FP32 links and arithmetic, FP64 reductions, FP64 measurement sweep) and,
as its check, in the FP64 mode.  At t in {1, 2, 4, 8} the full q(x) is
written in the reference's q(x) file format (float64 at the global index,
pi^2/4 scale, LGF/Observables/QdensityDistribution.cpp:15-28), read back,
and compared with the FP64 path; Q, E and t^2 E use FP64 accumulation.
At t = 1 the FP64 q(x) is spot-checked against the CPU oracle at a few
sites.  The reference itself cannot load or save a 64^4 configuration
(32-bit file offsets, LGF/Field/SU3_field.cpp:156,331), so this case has no
reference golden vector.

About 100 s of GPU time and ~60 GB of device memory in all, split into one
test per flow time (sharing the lattices through a module fixture) so that
each reports as it ends.
"""
import os

import numpy as np
import pytest

from msc_project_amd import lgf
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (64, 64, 64, 64)
BETA, EPS, SEED = 6.2, 0.01, 64646464
TIMES = (1.0, 2.0, 4.0, 8.0)


def site_q(shape, U, g):
    """per-site charge density of the oracle (improved F, the reference's
    plane pairs, LGF/Observables/TopologicalCharge.cpp:12-16), unscaled"""
    def herm(F):
        return np.asarray(F).reshape(9, 2)[:, 0] + 1j * np.asarray(F).reshape(9, 2)[:, 1]
    s = 0.0
    for a, b, c, d in ((3, 2, 0, 1), (3, 1, 2, 0), (3, 0, 1, 2)):
        Fa = herm(O.field_strength(shape, U, g, a, b, 4)).reshape(3, 3)
        Fb = herm(O.field_strength(shape, U, g, c, d, 4)).reshape(3, 3)
        s += np.real(np.trace(Fa @ Fb))
    return s


@pytest.fixture(scope="module")
def lattices():
    ctx = lgf.Context(0)
    ctx.autotune(SHAPE, lgf.FP32, 2)
    f64 = lgf.Field(ctx, SHAPE, lgf.FP64)
    f64.set_cold()
    f64.heatbath(BETA, SEED, sweeps=100, or_per_hb=4)
    ctx.synchronize()
    plaq = f64.diagnostics()[0]
    print("64^4 input: beta %.2f, 100 HB/OR iterations, <P> = %.6f" % (BETA, plaq), flush=True)
    f32 = lgf.Field(ctx, SHAPE, lgf.FP32)
    f32.copy_from(f64)
    state = {"ctx": ctx, "f64": f64, "f32": f32, "steps": 0, "plaq": plaq}
    yield state
    f32.close()
    f64.close()
    ctx.close()


def test_config5_input_thermalised(lattices):
    # <P> of beta 6.2 after 100 iterations from a cold start (literature: 0.6136 at equilibrium)
    assert 0.60 < lattices["plaq"] < 0.63


@pytest.mark.parametrize("t", TIMES)
def test_config5_qdensity_64(lattices, t, tmp_path):
    f64, f32 = lattices["f64"], lattices["f32"]
    n = int(round(t / EPS)) - lattices["steps"]
    assert n > 0
    f64.flow(EPS, n)
    f32.flow(EPS, n)
    lattices["steps"] += n
    o64 = f64.observables(lgf.MEASURE_DEFAULT, want_q=True)
    o32 = f32.observables(lgf.MEASURE_FP64, want_q=True)
    scale = np.pi ** 2 / 4  # q(x) file scale of the reference
    # q(x) files in the reference's format, read back
    paths = []
    for name, o in (("fp64", o64), ("fp32", o32)):
        p = str(tmp_path / ("torus_extdof1_0_Flowtime%f_%s.bin" % (t, name)))
        (o["qraw"] * scale).astype(np.float64).tofile(p)
        paths.append(p)
    q64 = np.fromfile(paths[0], dtype=np.float64)
    q32 = np.fromfile(paths[1], dtype=np.float64)
    assert q64.size == np.prod(SHAPE) and q32.size == np.prod(SHAPE)
    for p in paths:
        os.remove(p)
    Qabs = o64["Qabs"]
    # the file's sum is pi^4 Q (pi^2/4 scale times 4 pi^2 Q)
    assert abs(q64.sum() / (scale * 4 * np.pi ** 2) - o64["Q"]) < 1e-10 * Qabs
    dQ = abs(o32["Q"] - o64["Q"])
    dt2E = abs(o32["E"] - o64["E"]) / abs(o64["E"])
    dq = np.max(np.abs(q32 - q64)) / np.max(np.abs(q64))
    print("t = %.0f: Q64 %.6f Q32 %.6f |dQ| %.2e  t2E %.6f rel d(t2E) %.2e  max|dq|/max|q| %.2e" %
          (t, o64["Q"], o32["Q"], dQ, t * t * o64["E"], dt2E, dq), flush=True)
    assert dQ < 1e-4
    assert dt2E < 1e-6
    assert dq < 1e-3
    if t == TIMES[0]:
        # the FP64 path's per-site q(x) against the oracle, at sites spread over
        # the whole index range (the last one exercises the largest offsets)
        U = f64.download()
        V = np.prod(SHAPE)
        for g in (0, 12345, V // 2 + 777, V - 1):
            qo = site_q(SHAPE, U, g)
            assert abs(o64["qraw"][g] - qo) < 1e-12 * np.max(np.abs(o64["qraw"])), g
        del U
        print("t = 1: FP64 q(x) equals the oracle's at 4 sites", flush=True)
