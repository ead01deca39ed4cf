"""Shared helpers of the test suite (inputs, tolerances, golden files)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

# FP64 parity bar (relative) for E and Q per measurement against the reference
TOL_FP64 = 1e-12


def golden(name):
    with open(os.path.join(GOLDEN, "ref_%s.json" % name)) as f:
        return json.load(f)


def small_input():
    """The committed 4^4 fixture (hot start + 5 HB/OR sweeps, beta 6.0)."""
    d = golden("small")
    U = np.fromfile(os.path.join(GOLDEN, d["input_file"]), dtype=np.float64).reshape(-1, 4, 18)
    return d, U


def config1_inputs():
    """Regenerates the four 8^4 inputs of the config-1 golden case bit for bit."""
    from oracle import oracle as O
    d = golden("config1")
    g = d["generator"]
    shape = tuple(d["shape"])
    U = O.generate(shape, d["beta"], seed=g["seed"], start=g["start"], sweeps=g["therm_sweeps"],
                   or_per_hb=g["or_per_hb"])
    cfgs = [U.copy()]
    for k in range(3):
        O.generate(shape, d["beta"], seed=g["seed"], start="continue", sweeps=g["sep_sweeps"], U=U,
                   first_sweep=g["therm_sweeps"] + g["sep_sweeps"] * k, or_per_hb=g["or_per_hb"])
        cfgs.append(U.copy())
    return d, shape, cfgs


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def q_err(Q, Qref, Qabs):
    """|dQ| normalised by max(|Q|, sum_x |q(x)|): Q alone can be ~0."""
    Q, Qref, Qabs = (np.asarray(x, dtype=np.float64) for x in (Q, Qref, Qabs))
    return np.max(np.abs(Q - Qref) / np.maximum(np.abs(Qref), Qabs))


def haar_field(shape, seed):
    from oracle import oracle as O
    return O.generate(shape, 6.0, seed=seed, start="hot", sweeps=0)
