"""Serving loop: jobs that arrive while others flow are admitted by device
footprint, run batched, and each job's results are bitwise those of the same
job flowed alone, under either scheduling policy."""
import time

import numpy as np
import pytest

from msc_project_amd import lgf
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def alone(ctx, shape, eps, nsteps, every, prec, links=None, seed=None):
    f = lgf.Field(ctx, shape, prec)
    if links is not None:
        f.upload(links)
    else:
        f.set_random(seed)
    t, E, Q = [0.0], [], []
    e, q, _, _ = f.measure(lgf.MEASURE_FP64)
    E.append(e), Q.append(q)
    tt = 0.0
    for s in range(1, nsteps + 1):
        f.flow(eps, 1)
        tt += eps
        if s % every == 0 or s == nsteps:
            e, q, _, _ = f.measure(lgf.MEASURE_FP64)
            t.append(tt), E.append(e), Q.append(q)
    f.close()
    return np.array(t), np.array(E), np.array(Q)


@pytest.mark.parametrize("policy,min_sites", [(lgf.SERVE_SHORTEST_FIRST, 1 << 20), (lgf.SERVE_ALL, 0),
                                              (lgf.SERVE_SHORTEST_FIRST, 300)])
def test_server_results_same_alone_and_in_mix(policy, min_sites):
    jobs = [((4, 4, 4, 4), 0.02, 23, 5, lgf.FP32, 11), ((8, 4, 4, 6), 0.01, 17, 4, lgf.FP64, 12),
            ((6, 6, 4, 4), 0.03, 30, 10, lgf.FP32, 13), ((4, 4, 8, 8), 0.01, 9, 3, lgf.FP64, 14),
            ((4, 4, 4, 4), 0.02, 40, 7, lgf.FP32, 15)]
    srv = lgf.Server(0)
    srv.set_policy(policy, min_sites)  # 300 sites: only the shortest one or two jobs step at a time
    ids = []
    for k, (shape, eps, n, every, prec, seed) in enumerate(jobs):
        ids.append(srv.submit(shape, eps, n, every, precision=prec, seed=seed))
        time.sleep(0.002 * k)  # staggered arrivals: later jobs join a running batch
    res = [srv.wait(i, timeout=120) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["jobs_done"] == len(jobs)
    ctx = lgf.Context(0)
    for (shape, eps, n, every, prec, seed), r in zip(jobs, res):
        assert r["status"] == 0
        t, E, Q = alone(ctx, shape, eps, n, every, prec, seed=seed)
        assert np.array_equal(r["t"], t)
        assert np.array_equal(r["E"], E), (r["E"], E)
        assert np.array_equal(r["Q"], Q)
        assert r["submit_s"] <= r["start_s"] <= r["finish_s"]


def test_server_admission_by_footprint():
    shape = (8, 8, 8, 8)
    one = lgf.footprint(shape, lgf.FP32)
    reserve = 1 << 30  # the server keeps the measurement scratch (line products) aside
    srv = lgf.Server(0, budget_bytes=reserve + int(2.5 * one))  # two jobs fit at a time
    with pytest.raises(lgf.LgfError) as e:
        srv.submit((16, 16, 16, 16), 0.01, 5, 5)
    assert e.value.code == 8
    U = O.generate(shape, 6.0, seed=3, start="hot")
    ids = [srv.submit(shape, 0.01, 20, 10, links=U) for _ in range(5)]
    res = [srv.wait(i, timeout=120) for i in ids]
    srv.close()
    starts = sorted(r["start_s"] for r in res)
    finishes = sorted(r["finish_s"] for r in res)
    # the third job cannot start before the first finished
    assert starts[2] >= finishes[0] - 1e-6
    for r in res[1:]:
        assert np.array_equal(r["E"], res[0]["E"]) and np.array_equal(r["Q"], res[0]["Q"])


def test_server_autotune_only_when_idle():
    """The server's context is tuned only while no job is pending or
    resident; a job then runs with the tuned variant, bitwise as the same
    job flowed alone with that variant."""
    srv = lgf.Server(0)
    best = srv.autotune([(8, 8, 8, 8), (4, 4, 4, 8)], lgf.FP32, 2)
    assert 0 <= best < lgf.num_stage_variants(lgf.FP32)
    job = ((16, 16, 16, 16), 0.01, 300, 100, lgf.FP32, 21)
    shape, eps, n, every, prec, seed = job
    i = srv.submit(shape, eps, n, every, precision=prec, seed=seed)
    with pytest.raises(lgf.LgfError) as e:
        srv.autotune([(8, 8, 8, 8)], lgf.FP32, 2)
    assert e.value.code == 7  # LGF_E_STATE: a job is in flight
    r = srv.wait(i, timeout=120)
    assert r["status"] == 0 and len(r["t"]) == 4
    assert 0 <= srv.autotune([(16, 16, 16, 16)], lgf.FP32, 2)  # idle again
    srv.close()
    ctx = lgf.Context(0)
    ctx.set_stage_variant(lgf.FP32, best)
    t, E, Q = alone(ctx, shape, eps, n, every, prec, seed=seed)
    ctx.close()
    assert np.array_equal(r["t"], t) and np.array_equal(r["E"], E) and np.array_equal(r["Q"], Q)


def test_server_rejects_requests_beyond_result_capacity():
    """A result holds 1024 measurements; a request needing more is refused
    at submit (LGF_E_ARG), one needing exactly 1024 is accepted."""
    srv = lgf.Server(0)
    with pytest.raises(lgf.LgfError) as e:
        srv.submit((4, 4, 4, 4), 0.01, 2000, 1, seed=1)  # 2001 measurements
    assert e.value.code == 1
    with pytest.raises(lgf.LgfError) as e:
        srv.submit((4, 4, 4, 4), 0.01, 1024, 1, seed=1)  # t = 0 plus 1024
    assert e.value.code == 1
    i = srv.submit((4, 4, 4, 4), 0.001, 1023, 1, seed=1)  # exactly 1024
    r = srv.wait(i, timeout=120)
    srv.close()
    assert r["status"] == 0 and len(r["t"]) == 1024


def test_server_batch_creation_failure_fails_its_jobs():
    """When the scheduler cannot create the batch a job steps in, that job
    ends with the batch's status (LGF_E_NOMEM) after its t = 0 measurement,
    without stepping, its memory is freed, and later jobs run normally."""
    srv = lgf.Server(0)
    srv.inject_batch_failures(1)
    a = srv.submit((8, 8, 8, 8), 0.01, 20, 10, precision=lgf.FP32, seed=5)
    ra = srv.wait(a, timeout=120)
    assert ra["status"] == 4  # LGF_E_NOMEM
    assert np.array_equal(ra["t"], [0.0])
    b = srv.submit((8, 8, 8, 8), 0.01, 20, 10, precision=lgf.FP32, seed=5)
    rb = srv.wait(b, timeout=120)
    st = srv.stats()
    srv.close()
    assert rb["status"] == 0 and len(rb["t"]) == 3
    assert rb["E"][0] == ra["E"][0] and rb["Q"][0] == ra["Q"][0]
    assert st["jobs_done"] == 2
    ctx = lgf.Context(0)
    t, E, Q = alone(ctx, (8, 8, 8, 8), 0.01, 20, 10, lgf.FP32, seed=5)
    ctx.close()
    assert np.array_equal(rb["t"], t) and np.array_equal(rb["E"], E) and np.array_equal(rb["Q"], Q)


# shapes and couplings of the serving stream (the 500-request Poisson stream
# of the benchmark draws from these)
STREAM_SHAPES = [(16, 16, 16, 16), (24, 24, 24, 24), (32, 32, 32, 32), (32, 32, 32, 64)]
STREAM_BETAS = [5.96, 6.05, 6.13, 6.21]


def test_server_stream_shapes_same_alone_and_in_mix():
    """Jobs of the serving stream's lattice shapes (16^4 .. 32^3x64), short
    FP32 flows with mixed step sizes arriving staggered while others run:
    each job's E(t), Q(t) are bitwise those of the same job flowed alone, and
    its Q tracks the FP64 flow of the same input within |dQ| < 1e-4 and its
    E within 1e-6 relative."""
    rng = np.random.default_rng(20261204)
    ctx = lgf.Context(0)
    jobs = []
    for k in range(7):
        shape = STREAM_SHAPES[k % 4] if k < 4 else STREAM_SHAPES[rng.integers(4)]
        beta = STREAM_BETAS[rng.integers(4)]
        eps = (0.01, 0.02)[k % 2]
        nsteps = int(rng.integers(2, 5)) * 10
        g = lgf.Field(ctx, shape, lgf.FP64)
        g.set_cold()
        g.heatbath(beta, 7000 + k, sweeps=10, or_per_hb=4)
        jobs.append((shape, beta, eps, nsteps, g.download()))
        g.close()
    srv = lgf.Server(0)
    ids = []
    for k, (shape, beta, eps, nsteps, U) in enumerate(jobs):
        ids.append(srv.submit(shape, eps, nsteps, 10, precision=lgf.FP32, links=U, beta=beta))
        time.sleep(0.02)  # later jobs join a running batch
    res = [srv.wait(i, timeout=300) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["jobs_done"] == len(jobs)
    for (shape, beta, eps, nsteps, U), r in zip(jobs, res):
        assert r["status"] == 0
        assert len(r["t"]) == 1 + nsteps // 10
        t, E, Q = alone(ctx, shape, eps, nsteps, 10, lgf.FP32, links=U)
        assert np.array_equal(r["t"], t)
        assert np.array_equal(r["E"], E) and np.array_equal(r["Q"], Q), shape
        t64, E64, Q64 = alone(ctx, shape, eps, nsteps, 10, lgf.FP64, links=U)
        dQ = np.max(np.abs(r["Q"] - Q64))
        dE = np.max(np.abs(r["E"] - E64) / np.abs(E64))
        print("%s eps %.2f %d steps: max |dQ| %.2e, max rel dE %.2e vs FP64" % (shape, eps, nsteps, dQ, dE),
              flush=True)
        assert dQ < 1e-4 and dE < 1e-6
    ctx.close()
