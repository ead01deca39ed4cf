"""CPU checks of the C-ABI library: it loads without a GPU and exports every
function include/lgf_flow.h declares; host-side validation needs no GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from msc_project_amd import lgf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "lgf_flow.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lgf_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_api():
    names = declared_functions()
    assert len(names) >= 30
    assert set(names) == set(lgf.SIGNATURES), set(names) ^ set(lgf.SIGNATURES)


def test_binding_constants_match_the_header():
    src = open(os.path.join(ROOT, "include", "lgf_flow.h")).read()
    macros = dict(re.findall(r"#define (LGF_[A-Z0-9_]+) (.+)", src))
    assert lgf.FP32_SHARED_PLAQUETTE_VARIANTS == int(macros["LGF_FP32_SHARED_PLAQUETTE_VARIANTS"])
    assert lgf.FP32_COMPACT_VARIANTS == int(macros["LGF_FP32_COMPACT_VARIANTS"])
    assert lgf.FP64_SHARED_PLAQUETTE_VARIANTS == int(macros["LGF_FP64_SHARED_PLAQUETTE_VARIANTS"])
    orders = tuple(int(v) for v in macros["LGF_AUTOTUNE_TILE_ORDERS"].strip(" {}").split(","))
    assert lgf.AUTOTUNE_TILE_ORDERS == orders


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(lgf.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_footprint_and_shape_validation_without_gpu():
    # two link buffers (72 reals per site) and the accumulator (36 reals per site)
    assert lgf.footprint((4, 4, 4, 4), lgf.FP64) == (2 * 72 + 36) * 256 * 8
    assert lgf.footprint((24, 24, 24, 24), lgf.FP32) == (2 * 72 + 36) * 331776 * 4
    with pytest.raises(lgf.LgfError) as e:
        lgf.footprint((130, 128, 128, 128), lgf.FP32)  # beyond 2^28 sites
    assert e.value.code == 2
    with pytest.raises(lgf.LgfError) as e:
        lgf.footprint((4, 3, 4, 4), lgf.FP64)
    assert e.value.code == 2
    n = lgf.num_stage_variants(lgf.FP64)
    assert n >= 1 and lgf.num_stage_variants(lgf.FP32) >= 1


def test_context_without_gpu_fails_loudly():
    import subprocess
    import sys
    # a fresh process: no GPU in this container means lgf_ctx_create must error, never fall back
    code = "from msc_project_amd import lgf\ntry:\n    lgf.Context(0)\n    print('GPU')\nexcept lgf.LgfError as e:\n    print('ERR', e.code)\n"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert out.stdout.strip() in ("GPU", "ERR 5")


def test_flow_scale_t0_interpolates_t2E():
    """t0 from t^2 E = 0.3: linear interpolation between the bracketing points
    (checked against a closed form) and None when the series stays below"""
    t = np.linspace(0.0, 5.0, 101)
    E = 0.04 + 0.0 * t  # t^2 E = 0.04 t^2 crosses 0.3 at t = sqrt(7.5)
    t0 = lgf.flow_scale_t0(t, E)
    i = np.searchsorted(t ** 2 * E, 0.3)
    a, b = t[i - 1] ** 2 * E[i - 1] - 0.3, t[i] ** 2 * E[i] - 0.3
    assert abs(t0 - (t[i - 1] + (t[i] - t[i - 1]) * (-a) / (b - a))) < 1e-15
    assert abs(t0 - np.sqrt(7.5)) < 2e-3
    assert lgf.flow_scale_t0(t[:20], E[:20]) is None
    with pytest.raises(lgf.LgfError):
        lgf.flow_scale_t0(t[::-1], E)
