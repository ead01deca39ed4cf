"""ctypes binding of liblgf_flow.so (the C-ABI declared in include/lgf_flow.h).

The library is the only compute path: importing this module does not touch
the GPU, but every object that computes needs liblgf_flow.so and raises
LgfError when it is missing, so nothing can silently fall back to the CPU.
Link arrays are numpy float64 in the reference's saved-configuration order,
shape (V, 4, 18) (see include/lgf_flow.h).
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblgf_flow.so")

FP64 = 0
FP32 = 1
MEASURE_DEFAULT = 0
MEASURE_FP64 = 1
# mixed-precision stage variants below this one share each site's plaquettes
# between its four links; the others sum each link's staples in the FP64
# path's order (LGF_FP32_SHARED_PLAQUETTE_VARIANTS); from FP32_COMPACT_VARIANTS
# on the shared-plaquette variants also share the backward links and rebuild
# row 2 of shared links (LGF_FP32_COMPACT_VARIANTS)
FP32_SHARED_PLAQUETTE_VARIANTS = 8
FP32_COMPACT_VARIANTS = 6
# FP64 stage variants below this one share each site's plaquettes; the
# others sum each link's staples in the reference's order
# (LGF_FP64_SHARED_PLAQUETTE_VARIANTS)
FP64_SHARED_PLAQUETTE_VARIANTS = 3
# tile orders the autotuners time (LGF_AUTOTUNE_TILE_ORDERS)
AUTOTUNE_TILE_ORDERS = (0, 2, 4, 8, 404, 408, 604, 1204, 6402, 6404)
# variant ranges whose links agree bit for bit
FP32_VARIANT_GROUPS = ((0, FP32_COMPACT_VARIANTS), (FP32_COMPACT_VARIANTS, FP32_SHARED_PLAQUETTE_VARIANTS),
                       (FP32_SHARED_PLAQUETTE_VARIANTS, None))
FP64_VARIANT_GROUPS = ((0, FP64_SHARED_PLAQUETTE_VARIANTS), (FP64_SHARED_PLAQUETTE_VARIANTS, None))
SERVE_ALL = 0             # every admitted job steps in every round
SERVE_SHORTEST_FIRST = 1  # least remaining work first, at least min_sites per round
# lgf_observable order (include/lgf_flow.h)
OBSERVABLES = ("E", "Q", "Qabs", "E_improved", "Q_unimproved")

STATUS = {0: "ok", 1: "bad argument", 2: "bad shape", 3: "I/O error", 4: "out of memory", 5: "HIP error",
          6: "non-finite result", 7: "invalid state", 8: "admission refused", 9: "unsupported"}


class LgfError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__("lgf status %d (%s)%s" % (code, STATUS.get(code, "?"), (": " + msg) if msg else ""))


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)


class Request(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int * 4), ("precision", ctypes.c_int), ("beta", ctypes.c_double),
                ("eps", ctypes.c_double), ("nsteps", ctypes.c_int), ("measure_every", ctypes.c_int),
                ("links", _dp), ("start", ctypes.c_int), ("seed", ctypes.c_uint64), ("measure_flags", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("nmeas", ctypes.c_int),
                ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("finish_s", ctypes.c_double),
                ("t", ctypes.c_double * 1024), ("E", ctypes.c_double * 1024), ("Q", ctypes.c_double * 1024)]


# name -> (argtypes, restype); every symbol declared in include/lgf_flow.h
SIGNATURES = {
    "lgf_ctx_create": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "lgf_ctx_destroy": ([_vp], ctypes.c_int),
    "lgf_ctx_last_error": ([_vp], ctypes.c_char_p),
    "lgf_ctx_stream": ([_vp, ctypes.POINTER(_vp)], ctypes.c_int),
    "lgf_ctx_synchronize": ([_vp], ctypes.c_int),
    "lgf_ctx_set_stage_variant": ([_vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "lgf_ctx_get_stage_variant": ([_vp, ctypes.c_int, _ip], ctypes.c_int),
    "lgf_num_stage_variants": ([ctypes.c_int, _ip], ctypes.c_int),
    "lgf_autotune": ([_vp, _ip, ctypes.c_int, ctypes.c_int, _ip, _dp], ctypes.c_int),
    "lgf_autotune_batch": ([_vp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp], ctypes.c_int),
    "lgf_ctx_set_tile_order": ([_vp, ctypes.c_int], ctypes.c_int),
    "lgf_ctx_get_tile_order": ([_vp, _ip], ctypes.c_int),
    "lgf_ctx_set_batch_group_bytes": ([_vp, ctypes.c_size_t], ctypes.c_int),
    "lgf_ctx_get_batch_group_bytes": ([_vp, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "lgf_ctx_set_measure_scratch": ([_vp, ctypes.c_size_t], ctypes.c_int),
    "lgf_measure_observables": ([_vp, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "lgf_measure_observables_noalloc": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lgf_batch_measure_observables": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "lgf_batch_measure_start": ([_vp, ctypes.c_int], ctypes.c_int),
    "lgf_batch_measure_finish": ([_vp, _dp], ctypes.c_int),
    "lgf_measure_start": ([_vp, ctypes.c_int], ctypes.c_int),
    "lgf_measure_finish": ([_vp, _dp], ctypes.c_int),
    "lgf_field_footprint": ([_ip, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "lgf_field_create": ([_vp, _ip, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "lgf_field_destroy": ([_vp], ctypes.c_int),
    "lgf_field_shape": ([_vp, _ip, _ip], ctypes.c_int),
    "lgf_field_upload": ([_vp, _dp], ctypes.c_int),
    "lgf_field_download": ([_vp, _dp], ctypes.c_int),
    "lgf_field_set_cold": ([_vp], ctypes.c_int),
    "lgf_field_flow_time": ([_vp, _dp], ctypes.c_int),
    "lgf_field_set_flow_time": ([_vp, ctypes.c_double], ctypes.c_int),
    "lgf_batch_track_t0": ([_vp, ctypes.c_double], ctypes.c_int),
    "lgf_batch_t0": ([_vp, _dp], ctypes.c_int),
    "lgf_field_set_random": ([_vp, ctypes.c_uint64], ctypes.c_int),
    "lgf_field_heatbath": ([_vp, ctypes.c_double, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int],
                           ctypes.c_int),
    "lgf_field_copy": ([_vp, _vp], ctypes.c_int),
    "lgf_field_load_file": ([_vp, ctypes.c_char_p], ctypes.c_int),
    "lgf_field_save_file": ([_vp, ctypes.c_char_p], ctypes.c_int),
    "lgf_flow_steps": ([_vp, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "lgf_measure": ([_vp, ctypes.c_int, _dp, _dp, _dp, _dp], ctypes.c_int),
    "lgf_diagnostics": ([_vp, _dp, _dp, _dp], ctypes.c_int),
    "lgf_batch_create": ([_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "lgf_batch_destroy": ([_vp], ctypes.c_int),
    "lgf_batch_flow_steps": ([_vp, _dp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "lgf_batch_measure": ([_vp, ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
    "lgf_server_create": ([ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "lgf_server_destroy": ([_vp], ctypes.c_int),
    "lgf_server_submit": ([_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "lgf_server_wait": ([_vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(Result)], ctypes.c_int),
    "lgf_server_now": ([_vp, _dp], ctypes.c_int),
    "lgf_server_inject_batch_failures": ([_vp, ctypes.c_int], ctypes.c_int),
    "lgf_flow_scale_t0": ([_dp, _dp, ctypes.c_int, ctypes.c_double, _dp], ctypes.c_int),
    "lgf_server_set_policy": ([_vp, ctypes.c_int, ctypes.c_int64, ctypes.c_double], ctypes.c_int),
    "lgf_server_autotune": ([_vp, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip], ctypes.c_int),
    "lgf_server_stats": ([_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), _dp, _dp],
                         ctypes.c_int),
    "lgf_server_last_error": ([_vp], ctypes.c_char_p),
}


def lib():
    """Load liblgf_flow.so; raises LgfError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LgfError(9, "liblgf_flow.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _shape(shape):
    return (ctypes.c_int * 4)(*[int(s) for s in shape])


def volume(shape):
    v = 1
    for s in shape:
        v *= int(s)
    return v


_live = weakref.WeakSet()  # contexts and servers still open


@atexit.register
def _close_all():
    # release device objects in dependency order before the HIP runtime goes away
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        h = _vp()
        rc = lib().lgf_ctx_create(device, ctypes.byref(h))
        if rc:
            raise LgfError(rc, "lgf_ctx_create(%d)" % device)
        self.h = h
        self._fields = weakref.WeakSet()
        self._batches = weakref.WeakSet()
        _live.add(self)

    def check(self, rc, what=""):
        if rc:
            msg = lib().lgf_ctx_last_error(self.h)
            raise LgfError(rc, (what + ": " if what else "") + (msg.decode() if msg else ""))

    def stream(self):
        s = _vp()
        self.check(lib().lgf_ctx_stream(self.h, ctypes.byref(s)))
        return s.value

    def synchronize(self):
        self.check(lib().lgf_ctx_synchronize(self.h))

    def set_stage_variant(self, precision, variant):
        self.check(lib().lgf_ctx_set_stage_variant(self.h, precision, variant))

    def autotune(self, shape, precision, nsteps=10):
        """Times every stage-kernel variant; selects and returns (best, ms_per_step list)."""
        n = num_stage_variants(precision)
        ms = np.zeros(n)
        best = ctypes.c_int()
        self.check(lib().lgf_autotune(self.h, _shape(shape), precision, nsteps, ctypes.byref(best), _p(ms)),
                   "autotune")
        return best.value, ms.tolist()

    def autotune_batch(self, shapes, precision, nsteps=10):
        """Tunes tile order and stage variant on seeded lattices of `shapes`
        flowed as one batch; returns (best variant, ms_per_step list)."""
        n = num_stage_variants(precision)
        ms = np.zeros(n)
        best = ctypes.c_int()
        sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).reshape(-1, 4))
        self.check(lib().lgf_autotune_batch(self.h, sh.ctypes.data_as(_ip), len(sh), precision, nsteps,
                                            ctypes.byref(best), _p(ms)), "autotune_batch")
        return best.value, ms.tolist()

    def set_tile_order(self, block):
        self.check(lib().lgf_ctx_set_tile_order(self.h, int(block)))

    def set_measure_scratch(self, nbytes):
        self.check(lib().lgf_ctx_set_measure_scratch(self.h, int(nbytes)))

    def tile_order(self):
        v = ctypes.c_int()
        self.check(lib().lgf_ctx_get_tile_order(self.h, ctypes.byref(v)))
        return v.value

    def set_batch_group_bytes(self, nbytes):
        """cache-group cap of batches created afterwards (0: one group)"""
        self.check(lib().lgf_ctx_set_batch_group_bytes(self.h, int(nbytes)))

    def batch_group_bytes(self):
        v = ctypes.c_size_t()
        self.check(lib().lgf_ctx_get_batch_group_bytes(self.h, ctypes.byref(v)))
        return v.value

    def stage_variant(self, precision):
        v = ctypes.c_int()
        self.check(lib().lgf_ctx_get_stage_variant(self.h, precision, ctypes.byref(v)))
        return v.value

    def close(self):
        if self.h:
            for b in list(self._batches):
                b.close()
            for f in list(self._fields):
                f.close()
            lib().lgf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def num_stage_variants(precision):
    n = ctypes.c_int()
    rc = lib().lgf_num_stage_variants(precision, ctypes.byref(n))
    if rc:
        raise LgfError(rc)
    return n.value


def footprint(shape, precision):
    b = ctypes.c_size_t()
    rc = lib().lgf_field_footprint(_shape(shape), precision, ctypes.byref(b))
    if rc:
        raise LgfError(rc, "footprint")
    return b.value


class Field:
    """One lattice resident on the GPU."""

    def __init__(self, ctx, shape, precision=FP64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.precision = precision
        self.vol = volume(shape)
        h = _vp()
        ctx.check(lib().lgf_field_create(ctx.h, _shape(shape), precision, ctypes.byref(h)), "lgf_field_create")
        self.h = h
        self._batches = weakref.WeakSet()
        ctx._fields.add(self)

    def upload(self, links):
        links = np.ascontiguousarray(links, dtype=np.float64)
        assert links.size == self.vol * 72
        self.ctx.check(lib().lgf_field_upload(self.h, _p(links)), "upload")

    def download(self):
        out = np.empty((self.vol, 4, 18), dtype=np.float64)
        self.ctx.check(lib().lgf_field_download(self.h, _p(out)), "download")
        return out

    def set_cold(self):
        self.ctx.check(lib().lgf_field_set_cold(self.h), "set_cold")

    @property
    def flow_time(self):
        """t += eps per RK3 step taken (GradientFlow's flow time)"""
        t = ctypes.c_double()
        self.ctx.check(lib().lgf_field_flow_time(self.h, ctypes.byref(t)), "flow_time")
        return t.value

    @flow_time.setter
    def flow_time(self, t):
        self.ctx.check(lib().lgf_field_set_flow_time(self.h, float(t)), "set_flow_time")

    def set_random(self, seed):
        self.ctx.check(lib().lgf_field_set_random(self.h, int(seed)), "set_random")

    def heatbath(self, beta, seed, sweeps, or_per_hb=4, first_sweep=0):
        self.ctx.check(lib().lgf_field_heatbath(self.h, float(beta), int(seed), int(first_sweep), int(sweeps),
                                                int(or_per_hb)), "heatbath")

    def copy_from(self, other):
        self.ctx.check(lib().lgf_field_copy(self.h, other.h), "copy")

    def load_file(self, path):
        self.ctx.check(lib().lgf_field_load_file(self.h, path.encode()), "load_file")

    def save_file(self, path):
        self.ctx.check(lib().lgf_field_save_file(self.h, path.encode()), "save_file")

    def flow(self, eps, nsteps):
        self.ctx.check(lib().lgf_flow_steps(self.h, float(eps), int(nsteps)), "flow")

    def measure(self, flags=MEASURE_DEFAULT, want_q=False):
        """Returns (E, Q, Qabs, qraw or None)."""
        E, Q, Qa = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        q = np.empty(self.vol) if want_q else None
        self.ctx.check(lib().lgf_measure(self.h, flags, ctypes.byref(E), ctypes.byref(Q), ctypes.byref(Qa),
                                         _p(q) if want_q else None), "measure")
        return E.value, Q.value, Qa.value, q

    def observables(self, flags=MEASURE_DEFAULT, want_q=False):
        """All observables of one clover sweep as a dict (keys of OBSERVABLES),
        plus "qraw" (per-site density, or None)."""
        obs = np.empty(len(OBSERVABLES))
        q = np.empty(self.vol) if want_q else None
        self.ctx.check(lib().lgf_measure_observables(self.h, flags, _p(obs), _p(q) if want_q else None),
                       "measure_observables")
        out = dict(zip(OBSERVABLES, obs.tolist()))
        out["qraw"] = q
        return out

    def observables_noalloc(self, flags=MEASURE_DEFAULT):
        """observables() through the allocation-free entry point (no q(x)):
        its scratch is the library's static device array"""
        obs = np.empty(len(OBSERVABLES))
        self.ctx.check(lib().lgf_measure_observables_noalloc(self.h, flags, _p(obs)), "measure_observables_noalloc")
        return dict(zip(OBSERVABLES, obs.tolist()))

    def measure_start(self, flags=MEASURE_DEFAULT):
        """enqueue the clover sweep of the current links; collect with measure_finish"""
        self.ctx.check(lib().lgf_measure_start(self.h, flags), "measure_start")

    def measure_finish(self):
        """wait for the outstanding measurement; dict keyed by OBSERVABLES"""
        obs = np.empty(len(OBSERVABLES))
        self.ctx.check(lib().lgf_measure_finish(self.h, _p(obs)), "measure_finish")
        return dict(zip(OBSERVABLES, obs.tolist()))

    def diagnostics(self):
        p, mx, mn = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.ctx.check(lib().lgf_diagnostics(self.h, ctypes.byref(p), ctypes.byref(mx), ctypes.byref(mn)),
                       "diagnostics")
        return p.value, mx.value, mn.value

    def close(self):
        if self.h:
            for b in list(self._batches):
                b.close()
            if self.ctx.h:
                self.ctx.check(lib().lgf_field_destroy(self.h), "lgf_field_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flow_scale_t0(t, E, target=0.3):
    """t at which t^2 E(t) first reaches `target`, by linear interpolation of
    t^2 E (lgf_flow_scale_t0); None if the series never reaches it."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    E = np.ascontiguousarray(E, dtype=np.float64)
    out = ctypes.c_double()
    rc = lib().lgf_flow_scale_t0(_p(t), _p(E), len(t), float(target), ctypes.byref(out))
    if rc == 7:
        return None
    if rc:
        raise LgfError(rc, "flow_scale_t0")
    return out.value


class Batch:
    """Several lattices (any shapes, one precision) flowed by shared launches."""

    def __init__(self, ctx, fields):
        self.ctx = ctx
        self.fields = list(fields)
        arr = (_vp * len(self.fields))(*[f.h for f in self.fields])
        h = _vp()
        ctx.check(lib().lgf_batch_create(ctx.h, arr, len(self.fields), ctypes.byref(h)), "lgf_batch_create")
        self.h = h
        ctx._batches.add(self)
        for f in self.fields:
            f._batches.add(self)

    def flow(self, eps, nsteps, use_graph=False):
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float64), (len(self.fields),)))
        self.ctx.check(lib().lgf_batch_flow_steps(self.h, _p(e), int(nsteps), 1 if use_graph else 0), "batch flow")

    def measure(self, flags=MEASURE_DEFAULT):
        n = len(self.fields)
        E, Q, Qa = np.empty(n), np.empty(n), np.empty(n)
        self.ctx.check(lib().lgf_batch_measure(self.h, flags, _p(E), _p(Q), _p(Qa)), "batch measure")
        return E, Q, Qa

    def observables(self, flags=MEASURE_DEFAULT):
        """array (fields x len(OBSERVABLES)), columns in OBSERVABLES order"""
        obs = np.empty((len(self.fields), len(OBSERVABLES)))
        self.ctx.check(lib().lgf_batch_measure_observables(self.h, flags, _p(obs)), "batch measure_observables")
        return obs

    def measure_start(self, flags=MEASURE_DEFAULT):
        """enqueue the clover sweep of the current links; flow steps can be
        enqueued behind it before the result is collected with measure_finish"""
        self.ctx.check(lib().lgf_batch_measure_start(self.h, flags), "batch measure_start")

    def track_t0(self, target=0.3):
        """track t0 (t^2 E = target) on the device at every following measurement"""
        self.ctx.check(lib().lgf_batch_track_t0(self.h, float(target)), "batch track_t0")

    def t0(self):
        """t0 per field from the on-device tracker (NaN where not reached)"""
        out = np.empty(len(self.fields))
        self.ctx.check(lib().lgf_batch_t0(self.h, _p(out)), "batch t0")
        return out

    def measure_finish(self):
        """wait for the outstanding measurement: array (fields x len(OBSERVABLES))"""
        obs = np.empty((len(self.fields), len(OBSERVABLES)))
        self.ctx.check(lib().lgf_batch_measure_finish(self.h, _p(obs)), "batch measure_finish")
        return obs

    def close(self):
        if self.h:
            if self.ctx.h:
                lib().lgf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching flow server (lgf_server_*)."""

    def __init__(self, device=0, budget_bytes=0, max_active=0):
        h = _vp()
        rc = lib().lgf_server_create(device, budget_bytes, max_active, ctypes.byref(h))
        if rc:
            raise LgfError(rc, "lgf_server_create")
        self.h = h
        _live.add(self)

    def set_policy(self, policy, min_sites=1 << 20, age_rate=0.0):
        """age_rate: link-updates of remaining work credited per second waited"""
        self._check(lib().lgf_server_set_policy(self.h, int(policy), int(min_sites), float(age_rate)), "set_policy")

    def autotune(self, shapes, precision=FP32, nsteps=10):
        """Tunes the server's context for a mix of lattice shapes (idle server
        only); returns the chosen stage variant."""
        best = ctypes.c_int()
        sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).reshape(-1, 4))
        self._check(lib().lgf_server_autotune(self.h, sh.ctypes.data_as(_ip), len(sh), precision, nsteps,
                                              ctypes.byref(best)), "autotune")
        return best.value

    def _check(self, rc, what):
        if rc:
            msg = lib().lgf_server_last_error(self.h)
            raise LgfError(rc, what + ": " + (msg.decode() if msg else ""))

    def submit(self, shape, eps, nsteps, measure_every, precision=FP32, links=None, seed=None, beta=6.0,
               measure_flags=MEASURE_FP64):
        r = Request()
        r.shape[:] = [int(s) for s in shape]
        r.precision, r.beta, r.eps, r.nsteps, r.measure_every = precision, beta, eps, nsteps, measure_every
        r.measure_flags = measure_flags
        if links is not None:
            links = np.ascontiguousarray(links, dtype=np.float64)
            r.links = _p(links)
        else:
            r.links = None
            r.start = 1 if seed is not None else 0
            r.seed = int(seed or 0)
        i = ctypes.c_uint64()
        self._check(lib().lgf_server_submit(self.h, ctypes.byref(r), ctypes.byref(i)), "submit")
        return i.value

    def wait(self, job_id, timeout=-1.0):
        """Returns dict(status, t, E, Q, submit_s, start_s, finish_s)."""
        res = Result()
        self._check(lib().lgf_server_wait(self.h, job_id, float(timeout), ctypes.byref(res)), "wait")
        n = res.nmeas
        return {"id": res.id, "status": res.status, "t": np.array(res.t[:n]), "E": np.array(res.E[:n]),
                "Q": np.array(res.Q[:n]), "submit_s": res.submit_s, "start_s": res.start_s,
                "finish_s": res.finish_s}

    def inject_batch_failures(self, count):
        """test hook: the scheduler's next `count` batch creations fail (LGF_E_NOMEM)"""
        self._check(lib().lgf_server_inject_batch_failures(self.h, int(count)), "inject_batch_failures")

    def now(self):
        t = ctypes.c_double()
        self._check(lib().lgf_server_now(self.h, ctypes.byref(t)), "now")
        return t.value

    def stats(self):
        jd, bs = ctypes.c_uint64(), ctypes.c_uint64()
        lu, busy = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lgf_server_stats(self.h, ctypes.byref(jd), ctypes.byref(bs), ctypes.byref(lu),
                                           ctypes.byref(busy)), "stats")
        return {"jobs_done": jd.value, "batch_steps": bs.value, "link_updates": lu.value, "busy_s": busy.value}

    def close(self):
        if self.h:
            lib().lgf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
