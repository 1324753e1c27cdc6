"""Python face of the CPU oracle (test infrastructure only).

TEST INFRASTRUCTURE: only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this module, and only as the checker. It
mirrors the reference's Python API (python/main.cpp:46-226 for the shared
methods, python/main.cpp:475-1006 for the constructors and their defaults) on
top of ``_build/libfqsb_oracle.so`` (fqsb_oracle.c), so that a parity test can
build the same system from the oracle and from the product with the same
keyword arguments.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import types

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FQO_LIB: another build of fqsb_oracle.c (bench.py times a build for the
# host's own instruction set beside the portable parity build)
_LIB_PATH = os.environ.get("FQO_LIB") or os.path.join(_HERE, "_build", "libfqsb_oracle.so")

POT = {"Cuspy": 0, "SemiSmooth": 1, "Smooth": 2}
INT = {
    None: 0,
    "Laplace1d": 1,
    "Quartic1d": 2,
    "QuarticGradient1d": 3,
    "LongRange1d": 4,
    "Laplace2d": 5,
    "QuarticGradient2d": 6,
}
DIST = {
    "random": 0,
    "delta": 1,
    "exponential": 2,
    "power": 3,
    "gamma": 4,
    "pareto": 5,
    "weibull": 6,
    "normal": 7,
}
FIELD = {
    "u": 0,
    "v": 1,
    "a": 2,
    "f": 3,
    "f_potential": 4,
    "f_interactions": 5,
    "f_frame": 6,
    "f_damping": 7,
    "f_thermal": 8,
}


class _Desc(C.Structure):
    _fields_ = [
        ("rank", C.c_int32),
        ("potential", C.c_int32),
        ("interactions", C.c_int32),
        ("external", C.c_int32),
        ("minimisation", C.c_int32),
        ("distribution", C.c_int32),
        ("rows", C.c_uint64),
        ("cols", C.c_uint64),
        ("seed", C.c_uint64),
        ("nchunk", C.c_uint64),
        ("m", C.c_double),
        ("eta", C.c_double),
        ("mu", C.c_double),
        ("kappa", C.c_double),
        ("k_frame", C.c_double),
        ("dt", C.c_double),
        ("k_interactions", C.c_double),
        ("a1", C.c_double),
        ("a2", C.c_double),
        ("k2", C.c_double),
        ("k4", C.c_double),
        ("alpha", C.c_double),
        ("offset", C.c_double),
        ("parameters", C.c_double * 4),
        ("nparameters", C.c_int32),
        ("pad0", C.c_int32),
        ("mean", C.c_double),
        ("stddev", C.c_double),
        ("seed_forcing", C.c_uint64),
        ("dinc_init", C.POINTER(C.c_int64)),
        ("dinc", C.POINTER(C.c_int64)),
        ("align_buffer", C.c_int64),
        ("align_margin", C.c_int64),
    ]


_lib = None


def build():
    """Compile the oracle with its Makefile (a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int64)
    up = C.POINTER(C.c_uint64)
    sig = {
        "fqo_last_error": (C.c_char_p, []),
        "fqo_create": (C.c_int, [C.POINTER(_Desc), C.POINTER(P)]),
        "fqo_destroy": (None, [P]),
        "fqo_size": (C.c_uint64, [P]),
        "fqo_get": (C.c_int, [P, C.c_int, dp]),
        "fqo_set": (C.c_int, [P, C.c_int, dp]),
        "fqo_get_inc": (C.c_int64, [P]),
        "fqo_set_inc": (C.c_int, [P, C.c_int64]),
        "fqo_set_t": (C.c_int, [P, C.c_double]),
        "fqo_get_u_frame": (C.c_double, [P]),
        "fqo_set_u_frame": (C.c_int, [P, C.c_double]),
        "fqo_qs_first": (C.c_int64, [P]),
        "fqo_qs_last": (C.c_int64, [P]),
        "fqo_residual": (C.c_double, [P]),
        "fqo_temperature": (C.c_double, [P]),
        "fqo_refresh": (C.c_int, [P]),
        "fqo_quench": (C.c_int, [P]),
        "fqo_time_step": (C.c_int, [P]),
        "fqo_time_steps": (C.c_int, [P, C.c_uint64]),
        "fqo_flow_steps": (C.c_int, [P, C.c_uint64, C.c_double]),
        "fqo_time_steps_until_event": (C.c_int, [P, C.c_double, C.c_uint64, C.c_uint64, up]),
        "fqo_minimise": (C.c_int, [P, C.c_double, C.c_uint64, C.c_uint64, C.c_int, C.c_int, up]),
        "fqo_minimise_truncate": (
            C.c_int,
            [P, ip, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64, C.c_int, C.c_int, up],
        ),
        "fqo_max_uniform_displacement": (C.c_int, [P, C.c_int, dp]),
        "fqo_event_driven_step": (C.c_int, [P, C.c_double, C.c_int, C.c_int, dp]),
        "fqo_trigger": (C.c_int, [P, C.c_uint64, C.c_double, C.c_int]),
        "fqo_advance_to_fixed_force": (C.c_int, [P, C.c_double, C.c_int]),
        "fqo_chunk_size": (C.c_uint64, [P]),
        "fqo_chunk_data": (C.c_int, [P, dp]),
        "fqo_chunk_index_at_align": (C.c_int, [P, ip]),
        "fqo_chunk_chunk_index_at_align": (C.c_int, [P, ip]),
        "fqo_chunk_start": (C.c_int, [P, ip]),
        "fqo_chunk_state_at": (C.c_int, [P, ip, up]),
        "fqo_chunk_restore": (C.c_int, [P, up, dp, ip]),
        "fqo_thermal_state": (C.c_uint64, [P]),
        "fqo_thermal_set_state": (C.c_int, [P, C.c_uint64]),
        "fqo_thermal_next": (C.c_int, [P, ip]),
        "fqo_thermal_set_next": (C.c_int, [P, ip]),
        "fqo_pcg32_seed": (None, [C.c_uint64, C.c_uint64, up, up]),
        "fqo_pcg32_next": (C.c_uint32, [up, C.c_uint64]),
        "fqo_pcg32_advance": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_int64]),
        "fqo_uint32_to_double": (C.c_double, [C.c_uint32]),
        "fqo_draw": (C.c_int, [C.c_int, dp, C.c_int, up, C.c_uint64, dp, C.c_uint64]),
        "fqo_draw_normal": (C.c_int, [up, C.c_uint64, C.c_double, C.c_double, dp, C.c_uint64]),
        "fqo_pcg32_default_stream": (C.c_uint64, []),
        "fqo_log": (C.c_double, [C.c_double]),
        "fqo_exp": (C.c_double, [C.c_double]),
        "fqo_sin": (C.c_double, [C.c_double]),
        "fqo_cos": (C.c_double, [C.c_double]),
        "fqo_pow": (C.c_double, [C.c_double, C.c_double]),
        "fqo_stoplist_run": (C.c_int, [dp, C.c_uint64, C.c_uint64, C.c_double, ip]),
        "fqo_stoplist_margin": (C.c_int, [dp, C.c_uint64, C.c_uint64, C.c_double, dp]),
        "fqo_last_stop_margin": (C.c_double, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().fqo_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


# ---------------------------------------------------------------------------
# math / generator pins
# ---------------------------------------------------------------------------


def pcg32_seed(initstate, initseq):
    st = C.c_uint64()
    inc = C.c_uint64()
    lib().fqo_pcg32_seed(initstate, initseq, C.byref(st), C.byref(inc))
    return st.value, inc.value


def pcg32_next(state, inc, n):
    st = C.c_uint64(state)
    out = np.empty(n, dtype=np.uint32)
    for i in range(n):
        out[i] = lib().fqo_pcg32_next(C.byref(st), inc)
    return out, st.value


def pcg32_advance(state, inc, delta):
    return lib().fqo_pcg32_advance(state, inc, delta)


def draw(distribution, parameters, state, inc, n):
    par = np.zeros(4)
    par[: len(parameters)] = parameters
    st = C.c_uint64(state)
    out = np.empty(n)
    _check(lib().fqo_draw(DIST[distribution], _dp(par), len(parameters), C.byref(st), inc, _dp(out), n))
    return out, st.value


def draw_normal(state, inc, mean, stddev, n):
    """n consecutive thermal draws of one pcg32 stream (the draw of
    RandomNormalForcing::force, detail.h:935-940)."""
    st = C.c_uint64(state)
    out = np.empty(n)
    _check(lib().fqo_draw_normal(C.byref(st), inc, mean, stddev, _dp(out), n))
    return out, st.value


def forcing_stream(seed):
    """State and increment of the forcing generator of seed `seed`
    (prrng::pcg32::seed(seed) with the default stream, detail.h:917)."""
    return pcg32_seed(seed, lib().fqo_pcg32_default_stream())


def vec_math(name, x, y=None):
    fn = getattr(lib(), "fqo_" + name)
    x = np.asarray(x, dtype=float).ravel()
    if y is None:
        return np.array([fn(float(v)) for v in x])
    return np.array([fn(float(v), float(y)) for v in x])


def stoplist_margin(residuals, niter_tol, tol):
    """Decision margin of the StopList on a residual sequence: the smallest
    relative change of one residual that would change its stop step."""
    r = np.ascontiguousarray(residuals, dtype=float)
    m = C.c_double()
    _check(lib().fqo_stoplist_margin(_dp(r), r.size, niter_tol, tol, C.byref(m)))
    return m.value


# (system size, decision margin) of every minimise-type call of any system
# since the last clear; the GPU reduces the residual norms in another order
# than the oracle's sequential sum, so its stop step is the oracle's whenever
# the margin exceeds the rounding bound of the two sums.
STOP_MARGINS = []


def _record_margin(system):
    STOP_MARGINS.append((system.size, float(lib().fqo_last_stop_margin(system._h))))


def stoplist_run(residuals, niter_tol, tol):
    r = np.ascontiguousarray(residuals, dtype=float)
    idx = C.c_int64()
    _check(lib().fqo_stoplist_run(_dp(r), r.size, niter_tol, tol, C.byref(idx)))
    return idx.value


# ---------------------------------------------------------------------------
# system
# ---------------------------------------------------------------------------


class _Chunk:
    """Chunk of yield positions (the reference's prrng chunk, see Line1d.h:51)."""

    def __init__(self, sys):
        self._s = sys

    @property
    def data(self):
        s = self._s
        out = np.empty(s.size * s._nchunk)
        _check(lib().fqo_chunk_data(s._h, _dp(out)))
        return out.reshape(list(s.shape) + [s._nchunk])

    @property
    def index_at_align(self):
        out = np.empty(self._s.size, dtype=np.int64)
        _check(lib().fqo_chunk_index_at_align(self._s._h, _ip(out)))
        return out.reshape(self._s.shape)

    @property
    def chunk_index_at_align(self):
        out = np.empty(self._s.size, dtype=np.int64)
        _check(lib().fqo_chunk_chunk_index_at_align(self._s._h, _ip(out)))
        return out.reshape(self._s.shape)

    @property
    def start(self):
        out = np.empty(self._s.size, dtype=np.int64)
        _check(lib().fqo_chunk_start(self._s._h, _ip(out)))
        return out.reshape(self._s.shape)

    @property
    def left_of_align(self):
        d = self.data.reshape(self._s.size, -1)
        j = self.chunk_index_at_align.ravel()
        return d[np.arange(d.shape[0]), j].reshape(self._s.shape)

    @property
    def right_of_align(self):
        d = self.data.reshape(self._s.size, -1)
        j = self.chunk_index_at_align.ravel()
        return d[np.arange(d.shape[0]), j + 1].reshape(self._s.shape)

    def state_at(self, index):
        index = np.ascontiguousarray(np.broadcast_to(index, self._s.shape), dtype=np.int64).ravel()
        out = np.empty(self._s.size, dtype=np.uint64)
        _check(lib().fqo_chunk_state_at(self._s._h, _ip(index), _up(out)))
        return out.reshape(self._s.shape)

    def restore(self, state, value, index):
        sh = self._s.shape
        state = np.ascontiguousarray(np.broadcast_to(state, sh), dtype=np.uint64).ravel()
        value = np.ascontiguousarray(np.broadcast_to(value, sh), dtype=float).ravel()
        index = np.ascontiguousarray(np.broadcast_to(index, sh), dtype=np.int64).ravel()
        _check(lib().fqo_chunk_restore(self._s._h, _up(state), _dp(value), _ip(index)))


class _External:
    """RandomNormalForcing accessors (python/main.cpp:250-274)."""

    def __init__(self, sys):
        self._s = sys

    @property
    def state(self):
        return int(lib().fqo_thermal_state(self._s._h))

    @state.setter
    def state(self, value):
        _check(lib().fqo_thermal_set_state(self._s._h, int(value)))

    @property
    def f_thermal(self):
        return self._s._get("f_thermal")

    @f_thermal.setter
    def f_thermal(self, value):
        self._s._set("f_thermal", value)

    @property
    def next(self):
        out = np.empty(self._s.size, dtype=np.int64)
        _check(lib().fqo_thermal_next(self._s._h, _ip(out)))
        return out.reshape(self._s.shape)

    @next.setter
    def next(self, value):
        v = np.ascontiguousarray(np.broadcast_to(value, self._s.shape), dtype=np.int64).ravel()
        _check(lib().fqo_thermal_set_next(self._s._h, _ip(v)))


class _System:
    """Shared methods of every system (python/main.cpp:46-226)."""

    _has_dynamics = True
    _athermal = True

    def __init__(self, *, rank, shape, potential, interactions, external=0, minimisation=0, **kw):
        shape = [int(i) for i in shape]
        d = _Desc()
        d.rank = rank
        d.potential = POT[potential]
        d.interactions = INT[interactions]
        d.external = external
        d.minimisation = minimisation
        dist = kw["distribution"]
        if dist not in DIST:
            raise RuntimeError("Unknown distribution: " + dist)
        d.distribution = DIST[dist]
        d.rows = shape[0]
        d.cols = shape[1] if rank == 2 else 1
        d.seed = int(kw["seed"])
        d.nchunk = int(kw.get("nchunk", 5000))
        for name in ("m", "eta", "mu", "kappa", "k_frame", "dt", "k_interactions", "a1", "a2", "k2", "k4", "alpha"):
            setattr(d, name, float(kw.get(name, 0.0)))
        d.offset = float(kw.get("offset", -100.0))
        par = list(kw["parameters"])
        if len(par) > 4:
            raise RuntimeError("Too many parameters")
        for i, p in enumerate(par):
            d.parameters[i] = float(p)
        d.nparameters = len(par)
        d.align_buffer = 2
        d.align_margin = 30
        self._keep = []
        if external:
            d.mean = float(kw["mean"])
            d.stddev = float(kw["stddev"])
            d.seed_forcing = int(kw["seed_forcing"])
            n = int(np.prod(shape))
            di = np.ascontiguousarray(np.broadcast_to(kw["dinc_init"], [n]), dtype=np.int64)
            dd = np.ascontiguousarray(np.broadcast_to(kw["dinc"], [n]), dtype=np.int64)
            self._keep += [di, dd]
            d.dinc_init = _ip(di)
            d.dinc = _ip(dd)
        self._shape = tuple(shape)
        self._nchunk = d.nchunk
        self._dt = d.dt
        self._mu = d.mu
        self._eta = d.eta
        self._m = d.m
        self._k_frame = d.k_frame
        h = C.c_void_p()
        _check(lib().fqo_create(C.byref(d), C.byref(h)))
        self._h = h
        self._chunk = _Chunk(self)
        if external:
            self._external = _External(self)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.fqo_destroy(h)
            self._h = None

    # array helpers
    def _get(self, name):
        out = np.empty(self.size)
        _check(lib().fqo_get(self._h, FIELD[name], _dp(out)))
        return out.reshape(self._shape)

    def _set(self, name, value):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=float), self._shape)).ravel()
        _check(lib().fqo_set(self._h, FIELD[name], _dp(v)))

    # parameters
    size = property(lambda self: int(np.prod(self._shape)))
    shape = property(lambda self: list(self._shape))
    dt = property(lambda self: self._dt)
    mu = property(lambda self: self._mu)
    eta = property(lambda self: self._eta)
    m = property(lambda self: self._m)
    k_frame = property(lambda self: self._k_frame)
    chunk = property(lambda self: self._chunk)

    u = property(lambda self: self._get("u"), lambda self, v: self._set("u", v))
    v = property(lambda self: self._get("v"), lambda self, v: self._set("v", v))
    a = property(lambda self: self._get("a"), lambda self, v: self._set("a", v))
    f = property(lambda self: self._get("f"))
    f_potential = property(lambda self: self._get("f_potential"))
    f_frame = property(lambda self: self._get("f_frame"))
    f_interactions = property(lambda self: self._get("f_interactions"))
    f_damping = property(lambda self: self._get("f_damping"))

    @property
    def inc(self):
        return int(lib().fqo_get_inc(self._h))

    @inc.setter
    def inc(self, value):
        _check(lib().fqo_set_inc(self._h, int(value)))

    @property
    def t(self):
        return float(self.inc) * self._dt

    @t.setter
    def t(self, value):
        _check(lib().fqo_set_t(self._h, float(value)))

    @property
    def u_frame(self):
        return float(lib().fqo_get_u_frame(self._h))

    @u_frame.setter
    def u_frame(self, value):
        _check(lib().fqo_set_u_frame(self._h, float(value)))

    residual = property(lambda self: float(lib().fqo_residual(self._h)))
    temperature = property(lambda self: float(lib().fqo_temperature(self._h)))

    def refresh(self):
        _check(lib().fqo_refresh(self._h))

    def quench(self):
        _check(lib().fqo_quench(self._h))

    def maxUniformDisplacement(self, direction=1):
        r = C.c_double()
        _check(lib().fqo_max_uniform_displacement(self._h, int(direction), C.byref(r)))
        return r.value

    def trigger(self, p, eps, direction=1):
        _check(lib().fqo_trigger(self._h, int(p), float(eps), int(direction)))

    def advanceToFixedForce(self, f_frame, allow_plastic=False):
        _check(lib().fqo_advance_to_fixed_force(self._h, float(f_frame), int(bool(allow_plastic))))

    # athermal (python/main.cpp:158-202)
    def minimise(self, tol=1e-5, niter_tol=10, max_iter=int(1e9), time_activity=False, max_iter_is_error=True):
        r = C.c_uint64()
        _check(
            lib().fqo_minimise(
                self._h, tol, niter_tol, int(max_iter), int(time_activity), int(max_iter_is_error), C.byref(r)
            )
        )
        _record_margin(self)
        return r.value

    def minimise_truncate(
        self,
        i_n,
        A_truncate=0,
        S_truncate=0,
        tol=1e-5,
        niter_tol=10,
        max_iter=int(1e9),
        time_activity=True,
        max_iter_is_error=True,
    ):
        i_n = np.ascontiguousarray(np.broadcast_to(i_n, self._shape), dtype=np.int64).ravel()
        r = C.c_uint64()
        _check(
            lib().fqo_minimise_truncate(
                self._h,
                _ip(i_n),
                int(A_truncate),
                int(S_truncate),
                tol,
                niter_tol,
                int(max_iter),
                int(time_activity),
                int(max_iter_is_error),
                C.byref(r),
            )
        )
        _record_margin(self)
        return r.value

    def eventDrivenStep(self, eps, kick, direction=1):
        r = C.c_double()
        _check(lib().fqo_event_driven_step(self._h, float(eps), int(bool(kick)), int(direction), C.byref(r)))
        return r.value

    quasistaticActivityFirst = property(lambda self: int(lib().fqo_qs_first(self._h)))
    quasistaticActivityLast = property(lambda self: int(lib().fqo_qs_last(self._h)))

    # dynamics (python/main.cpp:210-226)
    def timeStep(self):
        _check(lib().fqo_time_step(self._h))

    def timeSteps(self, n):
        _check(lib().fqo_time_steps(self._h, int(n)))

    def timeStepsUntilEvent(self, tol=1e-5, niter_tol=10, max_iter=int(1e9)):
        r = C.c_uint64()
        _check(lib().fqo_time_steps_until_event(self._h, tol, niter_tol, int(max_iter), C.byref(r)))
        _record_margin(self)
        return r.value

    def flowSteps(self, n, v_frame):
        _check(lib().fqo_flow_steps(self._h, int(n), float(v_frame)))

    @property
    def external(self):
        return self._external


def _l1(potential, interactions, external=0, minimisation=0, **fixed):
    def make(cls_name):
        def __init__(self, **kw):
            kw = {**fixed, **kw}
            _System.__init__(
                self,
                rank=1,
                potential=potential,
                interactions=interactions,
                external=external,
                minimisation=minimisation,
                **kw,
            )

        return type(cls_name, (_System,), {"__init__": __init__})

    return make


def _l2(potential, interactions):
    def make(cls_name):
        def __init__(self, **kw):
            _System.__init__(self, rank=2, potential=potential, interactions=interactions, **kw)

        return type(cls_name, (_System,), {"__init__": __init__})

    return make


Line1d = types.SimpleNamespace(
    System_Cuspy_Laplace=_l1("Cuspy", "Laplace1d")("System_Cuspy_Laplace"),
    System_Cuspy_Laplace_Nopassing=_l1("Cuspy", "Laplace1d", minimisation=1, m=1.0)(
        "System_Cuspy_Laplace_Nopassing"
    ),
    System_Cuspy_Laplace_RandomForcing=_l1("Cuspy", "Laplace1d", external=1, minimisation=2)(
        "System_Cuspy_Laplace_RandomForcing"
    ),
    System_Cuspy_Quartic_RandomForcing=_l1("Cuspy", "Quartic1d", external=1, minimisation=2)(
        "System_Cuspy_Quartic_RandomForcing"
    ),
    System_SemiSmooth_Laplace=_l1("SemiSmooth", "Laplace1d")("System_SemiSmooth_Laplace"),
    System_Smooth_Laplace=_l1("Smooth", "Laplace1d")("System_Smooth_Laplace"),
    System_Cuspy_Quartic=_l1("Cuspy", "Quartic1d")("System_Cuspy_Quartic"),
    System_Cuspy_QuarticGradient=_l1("Cuspy", "QuarticGradient1d")("System_Cuspy_QuarticGradient"),
    System_Cuspy_LongRange=_l1("Cuspy", "LongRange1d")("System_Cuspy_LongRange"),
)

Particles = types.SimpleNamespace(
    System_Cuspy=_l1("Cuspy", None)("System_Cuspy"),
    System_Cuspy_RandomForcing=_l1("Cuspy", None, external=1, minimisation=2)("System_Cuspy_RandomForcing"),
    System_SemiSmooth=_l1("SemiSmooth", None)("System_SemiSmooth"),
    System_Smooth=_l1("Smooth", None)("System_Smooth"),
)

Line2d = types.SimpleNamespace(
    System_Cuspy_Laplace=_l2("Cuspy", "Laplace2d")("System_Cuspy_Laplace"),
    System_Cuspy_QuarticGradient=_l2("Cuspy", "QuarticGradient2d")("System_Cuspy_QuarticGradient"),
)
