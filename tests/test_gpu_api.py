"""Entry points and error paths of the C-ABI that the parity tests do not
reach: field shape and stream queries, the last-error text, the server
clock, and the status codes of misuse (empty batch, copy between shapes,
collecting a measurement that was never started, destroying a context that
still holds fields)."""
import ctypes

import numpy as np
import pytest

from msc_project_amd import lgf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = lgf.Context(0)
    yield c
    c.close()


def test_field_shape_and_stream(ctx):
    f = lgf.Field(ctx, (6, 4, 8, 2), lgf.FP32)
    shape, prec = (ctypes.c_int * 4)(), ctypes.c_int()
    assert lgf.lib().lgf_field_shape(f.h, shape, ctypes.byref(prec)) == 0
    assert list(shape) == [6, 4, 8, 2] and prec.value == lgf.FP32
    assert ctx.stream()  # a HIP stream handle
    f.close()


def test_status_codes_and_last_error(ctx):
    with pytest.raises(lgf.LgfError) as e:
        lgf.Field(ctx, (5, 4, 4, 4), lgf.FP64)
    assert e.value.code == 2  # LGF_E_SHAPE
    rc = lgf.lib().lgf_batch_create(ctx.h, None, 0, ctypes.byref(ctypes.c_void_p()))
    assert rc == 1  # LGF_E_ARG: a batch needs at least one field
    a, b = lgf.Field(ctx, (4, 4, 4, 4), lgf.FP64), lgf.Field(ctx, (4, 4, 4, 6), lgf.FP64)
    with pytest.raises(lgf.LgfError) as e:
        b.copy_from(a)
    assert e.value.code == 2
    assert lgf.lib().lgf_ctx_last_error(ctx.h)  # a message is left behind
    with pytest.raises(lgf.LgfError) as e:
        a.measure_finish()  # nothing outstanding
    assert e.value.code in (1, 7)
    # a context that still holds fields refuses to go, and changes nothing
    assert lgf.lib().lgf_ctx_destroy(ctx.h) == 7
    a.set_random(3)
    E, Q, _, _ = a.measure()
    assert np.isfinite(E) and np.isfinite(Q)
    a.close()
    b.close()


def test_server_clock_and_stats():
    srv = lgf.Server(0)
    t0 = srv.now()
    i = srv.submit((4, 4, 4, 4), 0.02, 10, 5, seed=1)
    r = srv.wait(i, timeout=60)
    t1 = srv.now()
    assert r["status"] == 0 and 0 <= t0 <= r["submit_s"] <= r["finish_s"] <= t1
    st = srv.stats()
    assert st["jobs_done"] == 1 and st["batch_steps"] >= 2 and st["link_updates"] == 4 * 256 * 10
    srv.close()
