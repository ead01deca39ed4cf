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
