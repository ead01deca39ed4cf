"""Flow-time scale t0 on the device, and the output files as the reference's
analysis scripts read them.

- lgf_batch_track_t0: the device-side t0 equals lgf_flow_scale_t0 on the
  host-collected (t, E) series, for synchronous and deferred measurements
  and for fields of one batch with different step sizes; the fields' flow
  times follow t += eps per step (LGF/Gradient_Flow/GradientFlow.cpp:52-58).
- The host program's CSVs read as DataAnalysis_and_Plotting/t0_extrapolation.py
  reads them (pandas.read_csv with index_col 't', names ['t', 'E<i>'], no
  header, lines 79-80, then a .loc window of flow times), and its q(x) file
  as QdensityDistribution.py reads it (numpy.fromfile float64, the first
  N2*N3 values reshaped to (N2, N3), lines 16-25): the values those readers
  see are the flow's, at the sites and flow times the scripts assume.
"""
import os

import numpy as np
import pytest

from msc_project_amd import lgf
from oracle import oracle as O
from tests.common import small_input
from tests.test_gpu_host import beta_dir, place_inputs, read_full, run_lgf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = lgf.Context(0)
    yield c
    c.close()


def thermalised(ctx, shape, beta, seed, sweeps=30):
    f = lgf.Field(ctx, shape, lgf.FP64)
    f.set_cold()
    f.heatbath(beta, seed, sweeps=sweeps, or_per_hb=4)
    return f


@pytest.mark.parametrize("deferred", [False, True])
def test_t0_on_device_matches_host_interpolation(ctx, deferred):
    shapes, epss = [(8, 8, 8, 8), (8, 8, 8, 12)], [0.02, 0.03]
    fields = []
    for k, shape in enumerate(shapes):
        g = thermalised(ctx, shape, 5.8, 100 + k)  # coarse: t0 ~ 1-2
        f = lgf.Field(ctx, shape, lgf.FP32)
        f.copy_from(g)
        g.close()
        fields.append(f)
    b = lgf.Batch(ctx, fields)
    b.track_t0(0.3)
    ts, Es = [[f.flow_time] for f in fields], []
    every, nmeas = 5, 40  # t up to 4 and 6, beyond t0 ~ 1-3 on these small, coarse lattices
    if deferred:
        b.measure_start()
        for m in range(nmeas):
            b.flow(epss, every, use_graph=True)
            Es.append(b.measure_finish()[:, 0])
            for i, f in enumerate(fields):
                ts[i].append(f.flow_time)
            if m < nmeas - 1:
                b.measure_start()
        Es = np.array(Es)
        ts = [t[:-1] for t in ts]
    else:
        Es.append(b.observables()[:, 0])
        for m in range(nmeas):
            b.flow(epss, every, use_graph=True)
            Es.append(b.observables()[:, 0])
            for i, f in enumerate(fields):
                ts[i].append(f.flow_time)
        Es = np.array(Es)
    t0_dev = b.t0()
    for i, eps in enumerate(epss):
        tt = 0.0
        want_t = [0.0]
        for s in range(1, every * nmeas + 1):
            tt += eps
            if s % every == 0:
                want_t.append(tt)
        if deferred:
            want_t = want_t[:-1]
        assert np.array_equal(ts[i], want_t)
        t0_host = lgf.flow_scale_t0(np.array(ts[i]), Es[:, i], 0.3)
        print("field %d: t0 device %.15g host %s" % (i, t0_dev[i], t0_host), flush=True)
        assert t0_host is not None
        assert abs(t0_dev[i] - t0_host) <= 1e-14 * t0_host
    b.close()
    for f in fields:
        f.close()


def test_t0_not_reached_is_nan_and_restart(ctx):
    g = thermalised(ctx, (8, 8, 8, 8), 6.0, 7)
    f = lgf.Field(ctx, (8, 8, 8, 8), lgf.FP64)
    f.copy_from(g)
    g.close()
    f.flow_time = 2.5  # as a resumed flow (SetFlowTime)
    b = lgf.Batch(ctx, [f])
    with pytest.raises(lgf.LgfError):
        b.t0()  # not tracking yet: LGF_E_STATE
    b.track_t0(100.0)  # t^2 E never reaches 100 in a few steps
    b.observables()
    b.flow(0.02, 5)
    b.observables()
    assert np.isnan(b.t0()[0])
    assert abs(f.flow_time - (2.5 + 5 * 0.02)) < 1e-15
    b.track_t0(1e-9)  # restarted: crossed between the next two measurements
    b.observables()
    b.flow(0.02, 1)
    b.observables()
    assert np.isnan(b.t0()[0])  # t^2 E was already above the target at the first measurement
    b.close()
    f.close()


def test_outputs_read_as_the_analysis_scripts_read_them(tmp_path):
    pd = pytest.importorskip("pandas")
    d, U = small_input()
    shape = tuple(d["shape"])
    folder = str(tmp_path) + "/"
    place_inputs(folder, shape, d["beta"], [U])
    run_lgf(folder, shape, d["beta"], 0, d["steps"], d["eps"], d["every"])
    run_lgf(folder, shape, d["beta"], 0, d["steps"], d["eps"], d["every"], "--mode=qdensity")
    sd = "X".join(map(str, shape))
    directory = os.path.join(folder, "Observables", "EnergyDensity", beta_dir(d["beta"]), sd, "GF") + "/"
    i = 0
    file = "torus_extdof4_" + str(i) + ".csv"
    df1 = pd.read_csv(directory + file, index_col="t", names=["t", "E" + str(i)], sep=",", header=None)
    _, E = read_full(os.path.join(directory, "torus_extdof4_0.full.csv"))
    assert len(df1) == len(E) == 1 + d["steps"] // d["every"]
    assert np.allclose(df1.index.values, np.arange(len(E)) * d["every"] * d["eps"], rtol=0, atol=1e-9)
    assert np.allclose(df1["E0"].values, E, rtol=1e-5, atol=0)  # the CSV's 6 significant digits
    window = df1.loc[0.15:0.35]  # a window of flow times, as the script selects t_comp_lim_low..high
    assert list(np.round(window.index.values, 6)) == [0.2, 0.3]
    # q(x): float64 at the global index; the script plots data[:xmax*ymax] as an (xmax, ymax) array,
    # the (x2, x3) plane at x0 = x1 = 0
    qx = os.path.join(folder, "Observables", "Qdensity", beta_dir(d["beta"]), sd, "GF",
                      "torus_extdof1_0_Flowtime%f.bin" % (d["steps"] * d["eps"]))
    with open(qx, "rb") as f:
        data = np.fromfile(f, dtype=np.float64)
    xmax, ymax = shape[2], shape[3]
    sliceArray = np.reshape(data[:xmax * ymax], (xmax, ymax))
    flowed = np.fromfile(os.path.join(folder, "ensembles", "su3", beta_dir(d["beta"]), sd, "GF",
                                      "torus_extdof4_0_Flowtime%f.bin" % (d["steps"] * d["eps"])),
                         dtype=np.float64).reshape(-1, 4, 18)
    from tests.test_gpu_large import site_q
    for x2 in range(xmax):
        for x3 in range(ymax):
            g = x2 * shape[3] + x3  # x0 = x1 = 0
            want = site_q(shape, flowed, g) * np.pi ** 2 / 4  # the file's scale (QdensityDistribution.cpp:15-28)
            assert abs(sliceArray[x2, x3] - want) < 1e-12 * np.max(np.abs(data)), (x2, x3)
