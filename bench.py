"""Benchmark of the Wilson-flow path on one MI355X.

Main workload (the single-lattice configuration of the headline metric): one
24^4 lattice at beta = 6.0, generated on the GPU from a seeded cold start by
200 heatbath/overrelaxation iterations (1 HB + 4 OR each), then flowed with
Luescher RK3, eps = 0.01, for 500 steps (t = 5), measuring E, t^2E and Q every
5 steps.  A "step" is one RK3 step of the whole lattice (4*24^4 link-updates)
with its share of the measurements.

`value`: link-updates/s of the whole job on the FP32-arithmetic path (FP32
links, staples, exponentials and intermediate clovers; FP64 reductions,
FP64 reunitarisation every step and an FP64 final measurement sweep), links
resident in HBM.  Reported beside it: the FP64 parity path on the same input,
flow-only rates, the accuracy of FP32 against FP64 at every measured flow
time and at t0 (t^2 E = 0.3), the roofline of the RK3 stage kernels (HIP
events on the library's stream), a batched run of 16 uneven lattices, a
continuous-batching serving run with p50/p99 latency, and a CPU baseline.

    python bench.py [--gpus 1] [--steps 500] [--warmup 10]
    python bench.py --profile fp32|fp32pb|fp32stream|fp64|fp32native --steps 6   (fixed workload for rocprofv3)
    python bench.py --profile fp32hold --steps 40000   (about 10 s of FP32 flow, for sampling the shader clock)
"""
import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SHAPE = (24, 24, 24, 24)
BETA = 6.0
EPS = 0.01
MEASURE_EVERY = 5
FLOW_TIMING_STEPS = 60  # flow-only timing: 180 stage launches between two HIP events
THERM_SWEEPS = 200
SEED = 24242424
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per link-update, the roofline basis: a link is read
# and written once per stage (6 transfers of 18 reals) and the RK3
# accumulator is written, read+written and read (4 transfers of a 3x3
# complex matrix), (6 + 4) * 18 * sizeof(real)
BYTES_PER_LU = {0: 1440.0, 1: 720.0}
# the same traffic with the accumulator stored as the 9 independent reals of
# a traceless anti-Hermitian matrix, as this library stores it:
# (6*18 + 4*9) * sizeof(real)
STORED_BYTES_PER_LU = {0: 1152.0, 1: 576.0}
ACCURACY_STEPS = 500  # the accuracy pass always flows to t = 5, past t0 (t^2 E = 0.3)


class HipEvents:
    """hipEvent timing on an arbitrary stream (the library's own)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def thermalised(lgf, ctx, shape, beta, seed, sweeps):
    f = lgf.Field(ctx, shape, lgf.FP64)
    f.set_cold()
    f.heatbath(beta, seed, sweeps)
    return f


def run_job(lgf, ctx, src, prec, steps, warmup, ev):
    """One flow job on a copy of `src`: returns timings and the measured series."""
    variant, tune_ms = ctx.autotune(SHAPE, prec, 10)
    flags = lgf.MEASURE_DEFAULT  # clover arithmetic in the path's precision, FP64 reductions
    # warm-up on a scratch copy (captures graphs, touches every kernel)
    w = lgf.Field(ctx, SHAPE, prec)
    w.copy_from(src)
    wb = lgf.Batch(ctx, [w])
    wb.flow(EPS, max(2, warmup), use_graph=True)
    wb.measure(flags)
    stream = ctx.stream()
    e0, e1 = ev.event(), ev.event()
    wb.flow(EPS, 10, use_graph=True)  # the replayed graphs, back to back
    ev.record(e0, stream)
    wb.flow(EPS, FLOW_TIMING_STEPS, use_graph=True)
    ev.record(e1, stream)
    flow_ms = ev.elapsed_ms(e0, e1) / FLOW_TIMING_STEPS
    ev.record(e0, stream)
    wb.measure(flags)
    ev.record(e1, stream)
    meas_ms = ev.elapsed_ms(e0, e1)
    wb.close()
    w.close()
    # the timed job
    f = lgf.Field(ctx, SHAPE, prec)
    f.copy_from(src)
    b = lgf.Batch(ctx, [f])
    ctx.synchronize()
    t0 = time.perf_counter()
    E, Q = [], []
    # deferred measurements (lgf_batch_measure_start / _finish): the next
    # MEASURE_EVERY flow steps are enqueued before the result is collected,
    # so the GPU does not wait for the host in between
    b.measure_start(flags)
    done = 0
    while done < steps:
        n = min(MEASURE_EVERY, steps - done)
        b.flow(EPS, n, use_graph=True)
        done += n
        o = b.measure_finish()[0]
        E.append(o[0]), Q.append(o[1])
        if done % MEASURE_EVERY == 0 and done < steps:
            b.measure_start(flags)
    e, q, _ = b.measure(lgf.MEASURE_FP64)  # final sweep in FP64
    E.append(e[0]), Q.append(q[0])
    ctx.synchronize()
    wall = time.perf_counter() - t0
    b.close()
    f.close()
    t = np.array([0.0] + [EPS * k for k in range(MEASURE_EVERY, steps + 1, MEASURE_EVERY)])[: len(E)]
    return {"wall_s": wall, "flow_ms_per_step": flow_ms, "measure_ms": meas_ms, "stage_variant": variant,
            "tile_order": ctx.tile_order(), "autotune_ms_per_step": tune_ms, "t": t, "E": np.array(E),
            "Q": np.array(Q)}


def accuracy_run(lgf, ctx, src, prec, steps=ACCURACY_STEPS):
    """Untimed flow of a copy of `src` for `steps` RK3 steps, E and Q every
    MEASURE_EVERY steps (final sweep in FP64, as in the timed job): the
    series the FP32-vs-FP64 accuracy is judged on, with the deviation from
    unitarity (max and mean of ||U^dag U - 1|| over the links) at every
    measurement."""
    flags = lgf.MEASURE_DEFAULT
    f = lgf.Field(ctx, SHAPE, prec)
    f.copy_from(src)
    b = lgf.Batch(ctx, [f])
    E, Q, unit = [], [], []

    def measure(fl):
        o = b.observables(fl)[0]
        E.append(o[0]), Q.append(o[1])
        unit.append(f.diagnostics()[1:])

    measure(flags)
    done = 0
    while done < steps:
        n = min(MEASURE_EVERY, steps - done)
        b.flow(EPS, n, use_graph=True)
        done += n
        measure(flags if done < steps else lgf.MEASURE_FP64)
    b.close()
    f.close()
    t = np.array([0.0] + [EPS * k for k in range(MEASURE_EVERY, steps + 1, MEASURE_EVERY)])
    unit = np.array(unit)
    return {"t": t, "E": np.array(E), "Q": np.array(Q), "unitarity": unit,
            "unitarity_max": float(unit[:, 0].max())}


def reference_order_fp64_rate(lgf, ctx, src, ev):
    """Flow-only rate of the FP64 per-link stage kernel, which sums each
    link's staples in the reference's order (tuned within its variant group,
    graph-replayed steps, HIP events on the library's stream).  The context's
    FP64 variant and tile order are restored afterwards."""
    keep_v, keep_o = ctx.stage_variant(lgf.FP64), ctx.tile_order()
    ctx.set_stage_variant(lgf.FP64, lgf.FP64_SHARED_PLAQUETTE_VARIANTS)
    variant, _ = ctx.autotune(SHAPE, lgf.FP64, 10)
    order = ctx.tile_order()
    f = lgf.Field(ctx, SHAPE, lgf.FP64)
    f.copy_from(src)
    b = lgf.Batch(ctx, [f])
    b.flow(EPS, 10, use_graph=True)
    stream = ctx.stream()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0, stream)
    b.flow(EPS, FLOW_TIMING_STEPS, use_graph=True)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1) / FLOW_TIMING_STEPS
    b.close()
    f.close()
    ctx.set_stage_variant(lgf.FP64, keep_v)
    ctx.set_tile_order(keep_o)
    return {"flow_only_lups": 4.0 * np.prod(SHAPE) / (ms * 1e-3), "flow_ms_per_step": ms, "stage_variant": variant,
            "tile_order": order}


def t0_of(t, E):
    from msc_project_amd import lgf
    return lgf.flow_scale_t0(t, E, 0.3)


def accuracy(r64, r32):
    t = r64["t"]
    y64, y32 = t ** 2 * r64["E"], t ** 2 * r32["E"]
    rel = np.abs(y32 - y64)[1:] / np.abs(y64[1:])
    dq = np.abs(r32["Q"] - r64["Q"])
    out = {"max_abs_dQ": float(dq.max()), "max_rel_dt2E": float(rel.max()), "measurements": int(len(t)),
           "final_abs_dQ": float(dq[-1]), "final_rel_dt2E": float(rel[-1]), "final_Q_fp64": float(r64["Q"][-1])}
    t0 = t0_of(t, r64["E"])
    if t0 is not None:
        i = int(np.searchsorted(t, t0))
        w = (t0 - t[i - 1]) / (t[i] - t[i - 1])

        def interp(a):
            return (1 - w) * a[i - 1] + w * a[i]

        out.update({"t0": float(t0), "abs_dQ_at_t0": float(interp(dq)),
                    "rel_dt2E_at_t0": float(abs(interp(y32) - interp(y64)) / interp(y64))})
    return out


def batched(lgf, ctx):
    """16 uneven lattices (6 x 12^4 at beta 5.8, 6 x 16^4 at 6.0, 4 x 20^3x40 at 6.1),
    each from 200 seeded heatbath/OR iterations, flowed together in FP32 to t = 3
    (300 steps of eps 0.01) with graph-captured steps, measuring every 10 steps."""
    specs = [((12, 12, 12, 12), 5.8)] * 6 + [((16, 16, 16, 16), 6.0)] * 6 + [((20, 20, 20, 40), 6.1)] * 4
    fields = []
    for i, (shape, beta) in enumerate(specs):
        g = thermalised(lgf, ctx, shape, beta, 1000 + i, THERM_SWEEPS)
        f = lgf.Field(ctx, shape, lgf.FP32)
        f.copy_from(g)
        g.close()
        fields.append(f)
    # tile order and stage variant tuned for this mix of shapes (seeded lattices, before the timed run)
    best, _ = ctx.autotune_batch([s for s, _ in specs], lgf.FP32, nsteps=5)
    lu = 4.0 * sum(np.prod(s) for s, _ in specs) * 300
    # the flow alone, on copies (graph-replayed steps)
    copies = []
    for f in fields:
        c = lgf.Field(ctx, f.shape, lgf.FP32)
        c.copy_from(f)
        copies.append(c)
    cb = lgf.Batch(ctx, copies)
    cb.flow(EPS, 4, use_graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    cb.flow(EPS, 60, use_graph=True)
    ctx.synchronize()
    flow_only = lu / 300 * 60 / (time.perf_counter() - t0)
    cb.close()
    for c in copies:
        c.close()
    b = lgf.Batch(ctx, fields)
    b.flow(EPS, 2, use_graph=True)  # capture
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(30):  # deferred measurements: the next flow steps are queued first
        b.flow(EPS, 10, use_graph=True)
        if k > 0:
            b.measure_finish()
        b.measure_start(lgf.MEASURE_DEFAULT)
    b.measure_finish()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    b.close()
    for f in fields:
        f.close()
    return {"configs": len(specs), "link_updates": lu, "wall_s": wall, "link_updates_per_s": lu / wall,
            "flow_only_link_updates_per_s": flow_only,
            "flow_only_hbm_roofline_frac": flow_only * BYTES_PER_LU[1] / (HBM_PEAK_GBS * 1e9),
            "roofline": {"bound": "hbm", "achieved": flow_only * BYTES_PER_LU[1] / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": flow_only * BYTES_PER_LU[1] / (HBM_PEAK_GBS * 1e9),
                         "frac_stored_bytes": flow_only * STORED_BYTES_PER_LU[1] / (HBM_PEAK_GBS * 1e9),
                         "basis": "720 B per link-update (FP32 links and accumulator); frac_stored_bytes: "
                                  "576 B, the accumulator as 9 reals",
                         "time": "wall clock of 60 graph-replayed RK3 steps of the whole batch"},
            "steps": 300, "measure_every": 10, "precision": "fp32",
            "tuning": {"stage_variant": best, "tile_order": ctx.tile_order(),
                       "cache_group_bytes": ctx.batch_group_bytes()}}


def serving(lgf, nreq, throughput, seed=4242, policy=1, age_rate=0.0):
    """Poisson stream of flow requests (shape uniform in {16^4, 24^4, 32^4,
    32^3x64}, beta in {5.96, 6.05, 6.13, 6.21}, eps 0.01, t uniform in [1, 6],
    measure every 10 steps, seeded random start) at a rate set for ~80%
    utilisation of the measured FP32 throughput; server policy `policy`
    (lgf.SERVE_SHORTEST_FIRST or lgf.SERVE_ALL)."""
    rng = np.random.default_rng(seed)
    shapes = [(16, 16, 16, 16), (24, 24, 24, 24), (32, 32, 32, 32), (32, 32, 32, 64)]
    betas = [5.96, 6.05, 6.13, 6.21]
    reqs = []
    for i in range(nreq):
        shape = shapes[rng.integers(4)]
        t = rng.uniform(1.0, 6.0)
        reqs.append((shape, betas[rng.integers(4)], int(round(t / EPS)), 100000 + i))
    mean_work = np.mean([4.0 * np.prod(s) * n for s, _, n, _ in reqs])
    rate = 0.8 * throughput / mean_work
    gaps = rng.exponential(1.0 / rate, size=nreq)
    srv = lgf.Server(0)
    srv.set_policy(policy, age_rate=age_rate)
    tuned = srv.autotune(shapes, lgf.FP32, nsteps=5)  # one lattice of each shape the stream draws from
    ids = [None] * nreq

    def feeder():
        t_next = time.perf_counter()
        for i, (shape, beta, n, sd) in enumerate(reqs):
            t_next += gaps[i]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids[i] = srv.submit(shape, EPS, n, 10, precision=lgf.FP32, seed=sd, beta=beta,
                                measure_flags=lgf.MEASURE_DEFAULT)
            if nreq >= 100 and (i + 1) % 50 == 0:
                log("serving: %d of %d requests submitted" % (i + 1, nreq))

    t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    res = []
    for k, i in enumerate(ids):
        res.append(srv.wait(i, timeout=600))
        if nreq >= 100 and (k + 1) % 50 == 0:
            log("serving: %d of %d requests done" % (k + 1, nreq))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([r["finish_s"] - r["submit_s"] for r in res])
    ok = all(r["status"] == 0 for r in res)
    lu = sum(4.0 * np.prod(s) * n for s, _, n, _ in reqs)
    return {"requests": nreq, "policy": "shortest_first" if policy == lgf.SERVE_SHORTEST_FIRST else "all",
            "age_rate": age_rate,
            "arrival_rate_per_s": rate, "all_ok": ok, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean()),
            "throughput_link_updates_per_s": lu / wall, "wall_s": wall, "busy_s": st["busy_s"],
            "batch_steps": st["batch_steps"], "stage_variant": tuned}


def stage_kernel_name(lgf, variant):
    """the mixed-precision stage kernel behind a stage variant"""
    if variant < lgf.FP32_COMPACT_VARIANTS:
        return "k_rk3_stage_pl<float>"
    if variant < lgf.FP32_SHARED_PLAQUETTE_VARIANTS:
        return "k_rk3_stage_pb<float>"
    return "k_rk3_stage<float>"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=20.0, threads=None):
    """The CPU oracle (a C port of the reference path, OpenMP over sites) on
    whole RK3 steps of a 24^4 lattice (seeded Haar-random links; the flow's
    CPU cost does not depend on the values), for about budget_s / 2 seconds."""
    from oracle import oracle as O
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    O.set_threads(threads)
    V = O.generate(SHAPE, BETA, seed=1234, start="hot", sweeps=0)
    t0 = time.perf_counter()
    n = 0
    while True:
        O.rk3_step(SHAPE, V, EPS)
        n += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    dt = time.perf_counter() - t0
    return {"value": 4.0 * np.prod(SHAPE) * n / dt, "unit": "link-updates/s", "cores": threads, "kind": "port",
            "host": {"cpu": cpu_model(), "logical_cpus": os.cpu_count()},
            "sample": "%d RK3 steps of a 24^4 lattice with the CPU oracle (oracle/lgf_oracle.c, %d OpenMP "
                      "threads), no measurements, %.1f s" % (n, threads, dt)}


def reference_record():
    """The reference program's own CPU throughput on a 24^4 lattice, from the
    committed record that oracle/ref/time_reference.py writes in the build
    container (the reference does not travel to the GPU host): flow-only and
    job rates at 1 and 8 MPI ranks, and the CPU oracle timed on that host, the
    bridge to this host's cpu_baseline."""
    p = os.path.join(ROOT, "profiles", "cpu_reference_flow_24x24x24x24.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    out = {"kind": "reference", "unit": "link-updates/s", "host": d["host"], "source": os.path.relpath(p, ROOT),
           "timed_where": "build container (the reference program is not run on the GPU host)"}
    for ranks, r in sorted(d["runs"].items(), key=lambda kv: int(kv[0])):
        out["ranks_%s" % ranks] = {"cores": r["cores"], "flow_only": r["flow_only_link_updates_per_s"],
                                   "job_measure_every_5": r["job_link_updates_per_s_measure_every_5"]}
    if "oracle_port" in d:
        out["oracle_port_same_host"] = d["oracle_port"]
    return out


STREAM_SHAPE = (32, 32, 32, 32)  # FP32 working set 604 MB: streams from HBM


def hold_flow(lgf, ctx, steps):
    """Sustained FP32 load for sampling the shader clock (rocm-smi): `steps`
    graph-replayed RK3 steps of a seeded random 24^4 lattice with the tuned
    stage kernel, in chunks of 100, printing the step time at the end."""
    ctx.autotune(SHAPE, lgf.FP32, 5)
    f = lgf.Field(ctx, SHAPE, lgf.FP32)
    f.set_random(5)
    b = lgf.Batch(ctx, [f])
    b.flow(0.001, 4, use_graph=True)
    ctx.synchronize()
    log("hold: start, variant %d, tile order %d" % (ctx.stage_variant(lgf.FP32), ctx.tile_order()))
    t0 = time.perf_counter()
    n = 0
    while n < steps:
        b.flow(0.001, 100, use_graph=True)
        ctx.synchronize()
        n += 100
    dt = time.perf_counter() - t0
    print(json.dumps({"profile": "fp32hold", "lattice": list(SHAPE), "steps": n, "seconds": dt,
                      "ms_per_step": dt / n * 1e3, "stage_variant": ctx.stage_variant(lgf.FP32)}))
    b.close()
    f.close()


def profile_run(which, steps, order=404):
    """Fixed workload for rocprofv3 counter runs: `steps` RK3 steps of a
    seeded random 24^4 lattice in one precision (direct launches, tile order
    `order`, x0/x2 blocks of 4 by default), two clover measurements, and one
    plain copy of a link array (k_convert, a kernel of known traffic that
    calibrates the counters).  "fp32stream" runs the same on a 32^4 lattice,
    which does not fit the last-level cache, with the shared-backward-link
    stage kernel with non-temporal stores (the autotuner's pick there)."""
    from msc_project_amd import lgf
    ctx = lgf.Context(0)
    if which == "fp32hold":
        hold_flow(lgf, ctx, steps)
        return
    prec = lgf.FP64 if which == "fp64" else lgf.FP32
    shape = STREAM_SHAPE if which == "fp32stream" else SHAPE
    ctx.set_tile_order(order)
    if which == "fp32pb":  # the shared-backward-link stage kernel
        ctx.set_stage_variant(lgf.FP32, lgf.FP32_COMPACT_VARIANTS)
    if which == "fp32stream":
        ctx.set_stage_variant(lgf.FP32, lgf.FP32_COMPACT_VARIANTS + 1)
    src = lgf.Field(ctx, shape, prec)  # copy source, cold by the time it is copied
    src.set_random(99)
    f = lgf.Field(ctx, shape, prec)
    f.set_random(1234)
    f.flow(EPS, steps)  # > 1 GB of traffic: evicts src from the 256 MB last-level cache
    for _ in range(2):  # the second sweep runs warm (instruction cache, TLB)
        f.measure(lgf.MEASURE_FP64 if which != "fp32native" else lgf.MEASURE_DEFAULT)
    g = lgf.Field(ctx, shape, prec)
    g.copy_from(src)
    ctx.synchronize()
    stride = (np.prod(shape) + 63) // 64 * 64
    nbytes = 72 * stride * (8 if prec == lgf.FP64 else 4)
    print(json.dumps({"profile": which, "lattice": list(shape), "steps": steps, "tile_order": order,
                      "copy_kernel_read_bytes": int(nbytes), "copy_kernel_write_bytes": int(nbytes),
                      "stage_algorithmic_bytes_per_step": BYTES_PER_LU[prec] * 4 * int(np.prod(shape))}))


def valu_floor(kernel, ms_per_step, sites):
    """VALU issue floor of an RK3 step of the named FP32 stage kernel, from
    the committed profiles: waves x VALU instructions per wave x 4 cycles
    over 1024 SIMDs at the shader clock measured under the same load, and
    its ratio to the measured step time."""
    p = os.path.join(ROOT, "profiles", "valu_floor_fp32_stage.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    ipw = d["valu_instr_per_wave"].get(kernel)
    if not ipw:
        return None
    waves = 4 * ((sites + 63) // 64)
    clock = d["sclk_ghz_under_load"]
    floor_ms = waves * sum(ipw) * 4 / (1024 * clock * 1e9) * 1e3
    return {"kernel": kernel, "clock_ghz": clock, "valu_instr_per_wave": ipw, "floor_ms_per_step": floor_ms,
            "achieved_ms_per_step": ms_per_step, "frac": floor_ms / ms_per_step,
            "source": "profiles/valu_floor_fp32_stage.json"}


def traffic_from_profiles(kernel):
    """HBM bytes per launch of the named FP32 stage kernel (mean of the three
    stages) from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary."""
    p = os.path.join(ROOT, "profiles", "rk3_stage_fp32_24x24x24x24_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch_by_kernel", {}).get(kernel)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="RK3 steps of the timed job")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=500,
                    help="requests of the serving stream (shortest-first policy); 0 disables it")
    ap.add_argument("--compare-requests", type=int, default=40,
                    help="requests of the comparison stream (every job steps); 0 disables it")
    ap.add_argument("--profile", choices=["fp32", "fp32pb", "fp32stream", "fp32hold", "fp64", "fp32native"], default=None,
                    help="run the fixed counter-profiling workload instead of the benchmark")
    ap.add_argument("--tile-order", type=int, default=404, help="tile order of the --profile workload")
    args = ap.parse_args()
    if args.profile:
        profile_run(args.profile, args.steps, args.tile_order)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the flow path runs one lattice (or one batch) per GPU "
              "and is not sharded across GPUs; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from msc_project_amd import lgf
    ev = HipEvents()
    ctx = lgf.Context(0)
    tg = time.perf_counter()
    src = thermalised(lgf, ctx, SHAPE, BETA, SEED, THERM_SWEEPS)
    ctx.synchronize()
    plaq = src.diagnostics()[0]
    log("input: 24^4 beta 6.0, %d HB/OR iterations in %.1f s, <P> = %.5f" % (THERM_SWEEPS, time.perf_counter() - tg,
                                                                             plaq))
    res, accr = {}, {}
    for prec in (lgf.FP64, lgf.FP32):
        name = "fp64" if prec == lgf.FP64 else "fp32"
        res[prec] = run_job(lgf, ctx, src, prec, args.steps, args.warmup, ev)
        r = res[prec]
        log("%s: job %.3f s, flow %.4f ms/step, measure %.3f ms, variant %d, tile order %d" % (
            name, r["wall_s"], r["flow_ms_per_step"], r["measure_ms"], r["stage_variant"], r["tile_order"]))
        # untimed accuracy pass to t = 5 with the same tuned kernels
        accr[prec] = accuracy_run(lgf, ctx, src, prec)
        log("%s: |U^dag U - 1| max/mean at t = 0, 0.05, ..., 5: %s" % (
            name, " ".join("%.2e/%.2e" % (a, m) for a, m in accr[prec]["unitarity"])))
    lu_step = 4.0 * np.prod(SHAPE)
    # beside the FP64 kernel the autotuner chose (shared plaquettes), the
    # per-link FP64 kernel that sums the staples in the reference's order
    ref64 = reference_order_fp64_rate(lgf, ctx, src, ev)
    log("fp64, per-link kernel (reference staple order): flow %.4f ms/step, variant %d, tile order %d" % (
        ref64["flow_ms_per_step"], ref64["stage_variant"], ref64["tile_order"]))
    r32, r64 = res[lgf.FP32], res[lgf.FP64]
    a32, a64 = accr[lgf.FP32], accr[lgf.FP64]
    acc = accuracy(a64, a32)
    acc["pass"] = "untimed flow of both precisions from the same input to t = %g (%d steps), E and Q every %d " \
                  "steps" % (ACCURACY_STEPS * EPS, ACCURACY_STEPS, MEASURE_EVERY)
    timed = accuracy(r64, r32)
    acc["timed_job"] = {k: timed[k] for k in ("max_abs_dQ", "max_rel_dt2E", "measurements")}
    log("accuracy: %s" % json.dumps(acc))
    value = lu_step * args.steps / r32["wall_s"]
    kern = stage_kernel_name(lgf, r32["stage_variant"])
    # one launch of a stage kernel advances every link of the lattice by one of
    # the three stages: 4V/3 link-updates; its mean duration from the HIP
    # events around FLOW_TIMING_STEPS graph-replayed steps on the library's stream
    launch_s = r32["flow_ms_per_step"] * 1e-3 / 3
    lu_launch = lu_step / 3
    achieved = BYTES_PER_LU[1] * lu_launch / launch_s / 1e9
    line = {
        "metric": "RK3 flow link-updates/s (FP32 arithmetic, FP64 reductions) on one 24^4 lattice",
        "value": value,
        "unit": "link-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall_s"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: 24^4 beta 6.0 from a seeded cold start and 200 GPU heatbath/OR iterations "
                "(<P> = %.5f)" % plaq,
        "config": {"workload": "single 24^4 lattice, beta 6.0, RK3 eps 0.01, E/t^2E/Q every 5 steps: %d steps timed "
                               "(t = %g); accuracy pass to t = %g" % (args.steps, args.steps * EPS,
                                                                      ACCURACY_STEPS * EPS),
                   "lattice": list(SHAPE), "eps": EPS, "measure_every": MEASURE_EVERY, "timed_steps": args.steps},
        "fp64_lups": lu_step * args.steps / r64["wall_s"],
        "fp32_over_fp64_job": r64["wall_s"] / r32["wall_s"],
        "fp32_flow_only_lups": lu_step / (r32["flow_ms_per_step"] * 1e-3),
        "fp64_flow_only_lups": lu_step / (r64["flow_ms_per_step"] * 1e-3),
        "fp32_over_fp64_flow": r64["flow_ms_per_step"] / r32["flow_ms_per_step"],
        "fp64_reference_order": dict(ref64, fp32_over=ref64["flow_ms_per_step"] / r32["flow_ms_per_step"],
                                     kernel="k_rk3_stage<double> (per link, the reference's staple order)"),
        "measure_ms": {"fp32": r32["measure_ms"], "fp64": r64["measure_ms"]},
        "fp32_vs_fp64": acc,
        "unitarity_max": {"fp32": a32["unitarity_max"], "fp64": a64["unitarity_max"]},
        "unitarity_every_measurement": {
            name: {"max_of_max": float(r["unitarity"][:, 0].max()), "max_of_mean": float(r["unitarity"][:, 1].max()),
                   "final_max": float(r["unitarity"][-1, 0]), "final_mean": float(r["unitarity"][-1, 1]),
                   "measurements": int(len(r["unitarity"]))} for name, r in (("fp32", a32), ("fp64", a64))},
        "tuning": {"stage_variant": {"fp64": r64["stage_variant"], "fp32": r32["stage_variant"]},
                   "tile_order": {"fp64": r64["tile_order"], "fp32": r32["tile_order"]}},
        "valu_floor": valu_floor(kern, r32["flow_ms_per_step"], int(np.prod(SHAPE))),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "frac_stored_bytes": STORED_BYTES_PER_LU[1] * lu_launch / launch_s / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles(kern),
                     "launch_us": launch_s * 1e6,
                     "kernel": "%s (stage variant %d, 3 launches per RK3 step)" % (kern, r32["stage_variant"]),
                     "basis": "720 B per link-update (FP32 links and accumulator) x 4V/3 link-updates per launch; "
                              "frac_stored_bytes: 576 B, the accumulator stored as its 9 independent reals"},
    }
    if not args.no_batch:
        line["batched_16_uneven"] = batched(lgf, ctx)
        log("batched: %s" % json.dumps(line["batched_16_uneven"]))
    if args.serve_requests > 0:
        # shortest remaining work first, with one second of waiting credited as
        # one second of the measured GPU throughput (bounds the wait of large jobs)
        line["serving"] = serving(lgf, args.serve_requests, value, policy=lgf.SERVE_SHORTEST_FIRST, age_rate=value)
        line["p99_latency_s"] = line["serving"]["p99_latency_s"]
        log("serving: %s" % json.dumps(line["serving"]))
    if args.compare_requests > 0:
        # a shorter stream of the same mix with every admitted job stepping each round
        line["serving_policy_all"] = serving(lgf, args.compare_requests, value, policy=lgf.SERVE_ALL)
        log("serving (all jobs step): %s" % json.dumps(line["serving_policy_all"]))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
        log("cpu baseline: %s" % json.dumps(line["cpu_baseline"]))
        ref = reference_record()
        if ref:
            line["cpu_baseline_reference"] = ref
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
