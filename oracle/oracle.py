"""ctypes front-end of the CPU restatement in lgf_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  Arrays are numpy
float64 in the reference's file order, shape (V, 4, 3, 3) complex128 viewed as
(V, 4, 18) float64 (see lgf_oracle.c for the layout and the reference lines
each function follows).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblgf_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liblgf_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_exp_ah.argtypes = [_dp, _dp]
        L.orc_staple.argtypes = [_ip, _dp, ctypes.c_long, ctypes.c_int, _dp]
        L.orc_flow_gradient.argtypes = [_ip, _dp, _dp]
        L.orc_rk3_step.argtypes = [_ip, _dp, ctypes.c_double]
        L.orc_field_strength.argtypes = [_ip, _dp, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_energy_density.argtypes = [_ip, _dp]
        L.orc_energy_density.restype = ctypes.c_double
        L.orc_top_charge.argtypes = [_ip, _dp, _dp, _dp]
        L.orc_top_charge.restype = ctypes.c_double
        L.orc_energy_density_kind.argtypes = [_ip, _dp, ctypes.c_int]
        L.orc_energy_density_kind.restype = ctypes.c_double
        L.orc_top_charge_kind.argtypes = [_ip, _dp, ctypes.c_int, _dp, _dp]
        L.orc_top_charge_kind.restype = ctypes.c_double
        L.orc_avg_plaquette.argtypes = [_ip, _dp]
        L.orc_avg_plaquette.restype = ctypes.c_double
        L.orc_unitarity.argtypes = [_ip, _dp, _dp, _dp]
        L.orc_flow_run.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_flow_run.restype = ctypes.c_int
        L.orc_generate.argtypes = [_ip, _dp, ctypes.c_double, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int]
        L.orc_reunitarize.argtypes = [_dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's site loops"""
    lib().orc_set_threads(int(n))


def _shape(shape):
    return (ctypes.c_int * 4)(*[int(s) for s in shape])


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def volume(shape):
    v = 1
    for s in shape:
        v *= int(s)
    return v


def empty_field(shape):
    return np.zeros((volume(shape), 4, 18), dtype=np.float64)


def exp_ah(X):
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(18)
    out = np.empty(18)
    lib().orc_exp_ah(_p(X), _p(out))
    return out


def staple(shape, U, g, mu):
    out = np.empty(18)
    lib().orc_staple(_shape(shape), _p(U), g, mu, _p(out))
    return out


def flow_gradient(shape, U):
    Z = np.empty_like(U)
    lib().orc_flow_gradient(_shape(shape), _p(U), _p(Z))
    return Z


def rk3_step(shape, U, eps):
    lib().orc_rk3_step(_shape(shape), _p(U), eps)


def field_strength(shape, U, g, mu, nu, kind):
    out = np.empty(18)
    lib().orc_field_strength(_shape(shape), _p(U), g, mu, nu, kind, _p(out))
    return out


def energy_density(shape, U, improved=False):
    """EnergyDensity, or ImprovedEnergyDensity when improved."""
    return lib().orc_energy_density_kind(_shape(shape), _p(U), 4 if improved else 0)


def top_charge_unimproved(shape, U):
    """UnImprovedTopologicalCharge: Q with the 1x1 clover F."""
    return lib().orc_top_charge_kind(_shape(shape), _p(U), 0, None, None)


def top_charge(shape, U, want_qx=False):
    """Returns (Q, sum_x |q(x)|, qx_file_values or None)."""
    qx = np.empty(volume(shape)) if want_qx else None
    a = ctypes.c_double(0.0)
    Q = lib().orc_top_charge(_shape(shape), _p(U), _p(qx) if want_qx else None, ctypes.byref(a))
    return Q, a.value, qx


def avg_plaquette(shape, U):
    return lib().orc_avg_plaquette(_shape(shape), _p(U))


def unitarity(shape, U):
    mx, mn = ctypes.c_double(), ctypes.c_double()
    lib().orc_unitarity(_shape(shape), _p(U), ctypes.byref(mx), ctypes.byref(mn))
    return mx.value, mn.value


def flow_run(shape, U, eps, nsteps, every):
    n = 1 + nsteps // every
    t, E, Q = np.empty(n), np.empty(n), np.empty(n)
    m = lib().orc_flow_run(_shape(shape), _p(U), eps, nsteps, every, _p(t), _p(E), _p(Q))
    return t[:m], E[:m], Q[:m]


def generate(shape, beta, seed, start="cold", sweeps=0, or_per_hb=4, U=None, first_sweep=0):
    """Seeded HB/OR chain. start in {"cold", "hot", "continue"}."""
    st = {"cold": 0, "hot": 1, "continue": -1}[start]
    if U is None:
        U = empty_field(shape)
    lib().orc_generate(_shape(shape), _p(U), beta, seed, st, first_sweep, sweeps, or_per_hb)
    return U


def reunitarize(m):
    lib().orc_reunitarize(_p(m))
