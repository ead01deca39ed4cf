"""The allocation-free measurement (lgf_measure_observables_noalloc, through
the C-ABI) against the allocating entry point it is a variant of
(lgf_measure_observables): bitwise equal observables on two calls in a row,
the second on a smaller lattice (the static scratch keeps the first call's
contents), in both precisions and both clover arithmetics; fields beyond the
scratch are refused with LGF_E_ARG."""
import numpy as np
import pytest

from msc_project_amd import lgf

pytestmark = pytest.mark.gpu


def field(ctx, shape, prec, seed):
    g = lgf.Field(ctx, shape, lgf.FP64)
    g.set_cold()
    g.heatbath(6.0, seed, sweeps=6, or_per_hb=2)
    f = lgf.Field(ctx, shape, prec)
    f.copy_from(g)
    g.close()
    f.flow(0.02, 3)
    return f


@pytest.mark.parametrize("prec,flags,shapes", [
    (lgf.FP32, lgf.MEASURE_DEFAULT, [(24, 24, 24, 24), (8, 8, 6, 4)]),
    (lgf.FP32, lgf.MEASURE_FP64, [(16, 16, 16, 16), (6, 4, 4, 8)]),
    (lgf.FP64, lgf.MEASURE_DEFAULT, [(20, 20, 20, 20), (12, 12, 12, 12)]),
])
def test_noalloc_matches_allocating_measurement(prec, flags, shapes):
    ctx = lgf.Context(0)
    ctx.set_tile_order(404)  # the allocating path's traversal order differs; results must not
    for k, shape in enumerate(shapes):  # the second call measures a smaller lattice
        f = field(ctx, shape, prec, 50 + k)
        want = f.observables(flags)
        got = f.observables_noalloc(flags)
        print(shape, prec, flags, got, flush=True)
        for key in lgf.OBSERVABLES:
            assert got[key] == want[key], (shape, key, got[key], want[key])
        f.close()
    ctx.close()


def test_noalloc_refuses_fields_beyond_its_scratch():
    ctx = lgf.Context(0)
    f = lgf.Field(ctx, (24, 24, 24, 24), lgf.FP64)  # 331776 sites > 232832 with FP64 arithmetic
    with pytest.raises(lgf.LgfError) as e:
        f.observables_noalloc()
    assert e.value.code == 1 and "232832" in str(e.value)
    f.close()
    g = lgf.Field(ctx, (32, 32, 32, 16), lgf.FP32)  # 524288 > 465408
    with pytest.raises(lgf.LgfError) as e:
        g.observables_noalloc()
    assert e.value.code == 1 and "465408" in str(e.value)
    g.close()
    ctx.close()
