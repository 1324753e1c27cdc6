"""Batched replicas: many systems of one description advanced in one launch.

No reference counterpart (the reference advances one system per object); this
is the batched form of the same operations, used for ensembles of replicas.
Replica r is built from landscape seed ``seed + r * seed_stride``; with
``seed_stride=1`` it equals the single system constructed with that seed
(Line1d.h:150: particle p of a system draws from stream ``seed + p``).

Thin ctypes layer over the C ABI of ``libfqsb.so`` (include/fqsb.h, the
``fqsb_batch_*`` functions); all work runs in the HIP library.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_here, "libfqsb.so")

POTENTIAL = {"Cuspy": 0, "SemiSmooth": 1, "Smooth": 2}
INTERACTIONS = {
    None: 0,
    "Laplace1d": 1,
    "Quartic1d": 2,
    "QuarticGradient1d": 3,
    "LongRange1d": 4,
    "Laplace2d": 5,
    "QuarticGradient2d": 6,
}
DISTRIBUTION = {
    "random": 0,
    "delta": 1,
    "exponential": 2,
    "power": 3,
    "gamma": 4,
    "pareto": 5,
    "weibull": 6,
    "normal": 7,
}
EXTERNAL = {None: 0, "RandomNormalForcing": 1}
MINIMISATION = {"Inertial": 0, "Overdamped": 1, "None": 2}
LONGRANGE_METHOD = {"exact": 0, "matrix": 1}
# checkpoint items (include/fqsb.h FQSB_STATE_*) -> (id, dtype)
STATE = {
    "y": (0, np.float64),
    "start": (1, np.int64),
    "jloc": (2, np.int32),
    "st_begin": (3, np.uint64),
    "st_end": (4, np.uint64),
    "f_thermal": (5, np.float64),
    "next": (6, np.int64),
    "inc": (7, np.int64),
    "u_frame": (8, np.float64),
    "rng_state": (9, np.uint64),
    "qs_first": (10, np.int64),
    "qs_last": (11, np.int64),
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


class Desc(C.Structure):
    """Mirror of ``fqsb_desc`` (include/fqsb.h)."""

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


_P = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
_up = C.POINTER(C.c_uint64)
_fp = C.POINTER(C.c_float)

# (restype, argtypes) of every function declared in include/fqsb.h
SIGNATURES = {
    "fqsb_last_error": (C.c_char_p, []),
    "fqsb_version": (C.c_char_p, []),
    "fqsb_create": (C.c_int, [C.POINTER(Desc), C.POINTER(_P)]),
    "fqsb_destroy": (C.c_int, [_P]),
    "fqsb_size": (C.c_uint64, [_P]),
    "fqsb_get": (C.c_int, [_P, C.c_int, _dp]),
    "fqsb_set": (C.c_int, [_P, C.c_int, _dp]),
    "fqsb_get_inc": (C.c_int, [_P, _ip]),
    "fqsb_set_inc": (C.c_int, [_P, C.c_int64]),
    "fqsb_set_t": (C.c_int, [_P, C.c_double]),
    "fqsb_get_u_frame": (C.c_int, [_P, _dp]),
    "fqsb_set_u_frame": (C.c_int, [_P, C.c_double]),
    "fqsb_quasistatic_activity": (C.c_int, [_P, _ip, _ip]),
    "fqsb_residual": (C.c_int, [_P, _dp]),
    "fqsb_temperature": (C.c_int, [_P, _dp]),
    "fqsb_refresh": (C.c_int, [_P]),
    "fqsb_quench": (C.c_int, [_P]),
    "fqsb_time_step": (C.c_int, [_P]),
    "fqsb_time_steps": (C.c_int, [_P, C.c_uint64]),
    "fqsb_flow_steps": (C.c_int, [_P, C.c_uint64, C.c_double]),
    "fqsb_time_steps_until_event": (C.c_int, [_P, C.c_double, C.c_uint64, C.c_uint64, _up]),
    "fqsb_minimise": (C.c_int, [_P, C.c_double, C.c_uint64, C.c_uint64, C.c_int, C.c_int, _up]),
    "fqsb_minimise_truncate": (
        C.c_int,
        [_P, _ip, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64, C.c_int, C.c_int, _up],
    ),
    "fqsb_max_uniform_displacement": (C.c_int, [_P, C.c_int, _dp]),
    "fqsb_event_driven_step": (C.c_int, [_P, C.c_double, C.c_int, C.c_int, _dp]),
    "fqsb_trigger": (C.c_int, [_P, C.c_uint64, C.c_double, C.c_int]),
    "fqsb_advance_to_fixed_force": (C.c_int, [_P, C.c_double, C.c_int]),
    "fqsb_chunk_size": (C.c_uint64, [_P]),
    "fqsb_chunk_data": (C.c_int, [_P, _dp]),
    "fqsb_chunk_index_at_align": (C.c_int, [_P, _ip]),
    "fqsb_chunk_chunk_index_at_align": (C.c_int, [_P, _ip]),
    "fqsb_chunk_start": (C.c_int, [_P, _ip]),
    "fqsb_chunk_state_at": (C.c_int, [_P, _ip, _up]),
    "fqsb_chunk_restore": (C.c_int, [_P, _up, _dp, _ip]),
    "fqsb_thermal_state": (C.c_int, [_P, _up]),
    "fqsb_thermal_set_state": (C.c_int, [_P, C.c_uint64]),
    "fqsb_thermal_next": (C.c_int, [_P, _ip]),
    "fqsb_thermal_set_next": (C.c_int, [_P, _ip]),
    "fqsb_batch_create": (C.c_int, [C.POINTER(Desc), C.c_uint64, C.c_uint64, C.POINTER(_P)]),
    "fqsb_batch_destroy": (C.c_int, [_P]),
    "fqsb_batch_replicas": (C.c_uint64, [_P]),
    "fqsb_batch_set_launch": (C.c_int, [_P, C.c_int, C.c_int]),
    "fqsb_batch_get_launch": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fqsb_batch_set_longrange_method": (C.c_int, [_P, C.c_int]),
    "fqsb_batch_get_longrange_method": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "fqsb_batch_longrange_product_ms": (C.c_int, [_P, _fp]),
    "fqsb_batch_flow_steps": (C.c_int, [_P, C.c_uint64, C.c_double, C.c_uint64, _fp]),
    "fqsb_batch_time_steps": (C.c_int, [_P, C.c_uint64, C.c_uint64, _fp]),
    "fqsb_batch_minimise": (C.c_int, [_P, C.c_double, C.c_uint64, C.c_uint64, C.c_int, _up]),
    "fqsb_batch_event_driven_step": (C.c_int, [_P, C.c_double, C.c_int, C.c_int, _dp]),
    "fqsb_batch_quasistatic": (
        C.c_int,
        [_P, C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint64, _ip, _dp, _dp,
         _ip],
    ),
    "fqsb_batch_quasistatic_noalloc": (
        C.c_int,
        [_P, C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint64, _ip, _dp, _dp,
         _ip],
    ),
    "fqsb_batch_quasistatic_stats": (C.c_int, [_P, _fp, _ip]),
    "fqsb_batch_minimise_truncate": (
        C.c_int,
        [_P, _ip, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64, C.c_int, _up],
    ),
    "fqsb_batch_trigger": (C.c_int, [_P, _up, C.c_double, C.c_int]),
    "fqsb_batch_state_size": (C.c_int, [_P, C.c_int, _up, C.POINTER(C.c_int)]),
    "fqsb_batch_get_state": (C.c_int, [_P, C.c_int, C.c_void_p]),
    "fqsb_batch_set_state": (C.c_int, [_P, C.c_int, C.c_void_p]),
    "fqsb_batch_refresh": (C.c_int, [_P]),
    "fqsb_batch_advance_to_fixed_force": (C.c_int, [_P, _dp, C.c_int]),
    "fqsb_quasistatic": (
        C.c_int,
        [_P, C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint64, _ip, _dp, _dp,
         _ip],
    ),
    "fqsb_batch_get": (C.c_int, [_P, C.c_int, _dp]),
    "fqsb_batch_set": (C.c_int, [_P, C.c_int, _dp]),
    "fqsb_batch_get_u_frame": (C.c_int, [_P, _dp]),
    "fqsb_batch_get_inc": (C.c_int, [_P, _ip]),
    "fqsb_batch_chunk_index_at_align": (C.c_int, [_P, _ip]),
    "fqsb_batch_stream": (C.c_void_p, [_P]),
    "fqsb_batch_set_parameters": (C.c_int, [_P, _dp]),
    "fqsb_batch_get_parameters": (C.c_int, [_P, _dp]),
    "fqsb_batch_nchunk": (C.c_uint64, [_P]),
    "fqsb_batch_get_device": (C.c_int, [_P, C.c_int, C.c_void_p]),
    "fqsb_batch_set_device": (C.c_int, [_P, C.c_int, C.c_void_p]),
    "fqsb_batch_chunk_data_device": (C.c_int, [_P, C.c_void_p]),
    "fqsb_batch_set_chunk_data_device": (C.c_int, [_P, C.c_void_p]),
    "fqsb_batch_adjoint_forward": (C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_int, C.c_double]),
    "fqsb_batch_adjoint_backward_device": (C.c_int, [_P] + [C.c_void_p] * 8),
    "fqsb_batch_adjoint_kernel_ms": (C.c_int, [_P, _fp, _fp]),
    "fqsb_batch_adjoint_generation": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "fqsb_ragged_create": (C.c_int, [C.POINTER(Desc), C.c_uint64, C.POINTER(_P)]),
    "fqsb_ragged_destroy": (C.c_int, [_P]),
    "fqsb_ragged_systems": (C.c_uint64, [_P]),
    "fqsb_ragged_groups": (C.c_uint64, [_P]),
    "fqsb_ragged_sizes": (C.c_int, [_P, _up]),
    "fqsb_ragged_time_steps": (C.c_int, [_P, C.c_uint64]),
    "fqsb_ragged_flow_steps": (C.c_int, [_P, C.c_uint64, C.c_double]),
    "fqsb_ragged_minimise": (
        C.c_int,
        [_P, C.c_double, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_uint8), _up],
    ),
    "fqsb_ragged_event_driven_step": (
        C.c_int,
        [_P, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_uint8), _dp],
    ),
    "fqsb_ragged_quasistatic": (
        C.c_int,
        [_P, C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint64,
         C.POINTER(C.c_uint8), _ip, _dp, _dp, _ip],
    ),
    "fqsb_ragged_get": (C.c_int, [_P, C.c_int, _dp]),
    "fqsb_ragged_set": (C.c_int, [_P, C.c_int, _dp]),
    "fqsb_ragged_get_u_frame": (C.c_int, [_P, _dp]),
    "fqsb_ragged_get_inc": (C.c_int, [_P, _ip]),
    "fqsb_ragged_last_steps": (C.c_int, [_P, _up]),
    "fqsb_ragged_chunk_index_at_align": (C.c_int, [_P, _ip]),
    "fqsb_ragged_group_info": (C.c_int, [_P, C.c_uint64, _ip]),
    "fqsb_ragged_group_systems": (C.c_int, [_P, C.c_uint64, _up]),
    "fqsb_ragged_group_time": (C.c_int, [_P, C.c_uint64, _dp, _dp]),
    "fqsb_ragged_reset_group_time": (C.c_int, [_P]),
    "fqsb_ragged_set_serial": (C.c_int, [_P, C.c_int]),
    "fqsb_ragged_group_shapes": (
        C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
    ),
    "fqsb_ragged_set_group_launch": (C.c_int, [_P, C.c_uint64, C.c_int, C.c_int]),
    "fqsb_ragged_set_group_block": (C.c_int, [_P, C.c_uint64, C.c_int]),
}

_lib = None


def capi():
    """The C ABI of libfqsb.so with argument types declared."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(f"{LIBRARY_PATH} is not built; run `python -m frictionqpotspringblock_amd.build_ext`")
        L = C.CDLL(LIBRARY_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _quasistatic(call, nsys, eps, nevents, first_event, direction, kicks_only, tol, niter_tol, max_iter, masked=False,
                 mask=None):
    """Shared wrapper of fqsb_*_quasistatic: outputs [nsys, nevents]; `masked`:
    the entry point takes an `active` mask (None: every system)."""
    nevents = int(nevents)
    n = max(nevents, 0)
    out = dict(
        S=np.zeros([nsys, n], dtype=np.int64),
        u_frame=np.zeros([nsys, n]),
        f_frame=np.zeros([nsys, n]),
        inc=np.zeros([nsys, n], dtype=np.int64),
    )
    args = [int(first_event), nevents, float(eps), int(direction), int(bool(kicks_only)), float(tol), int(niter_tol),
            int(max_iter)]
    if masked:
        args.append(None if mask is None else mask.ctypes.data_as(C.POINTER(C.c_uint8)))
    args += [
        out["S"].ctypes.data_as(_ip),
        out["u_frame"].ctypes.data_as(_dp),
        out["f_frame"].ctypes.data_as(_dp),
        out["inc"].ctypes.data_as(_ip),
    ]
    _check(call(*args))
    return out


def validate_checkpoint(ck, replicas, shape, nchunk, external):
    """Check a checkpoint (Batch.checkpoint, or an .npz written by
    Batch.save) against a batch of `replicas` systems of `shape` with windows
    of `nchunk` yield positions: every item present with its size, and every
    local well index inside the window (the device reads y[p][jloc] and
    y[p][jloc + 1]). Raises RuntimeError; runs on the host only."""
    n = replicas * int(np.prod(shape))
    sizes = {
        "u": n, "v": n, "a": n, "y": n * nchunk, "start": n, "jloc": n, "st_begin": n, "st_end": n,
        "inc": replicas, "u_frame": replicas, "qs_first": replicas, "qs_last": replicas,
    }
    if external:
        sizes.update(f_thermal=n, next=n, rng_state=replicas)
    for k, size in sizes.items():
        if k not in ck:
            raise RuntimeError(f"checkpoint item {k} is missing")
        if np.size(ck[k]) != size:
            raise RuntimeError(f"checkpoint item {k} has {np.size(ck[k])} values, the batch needs {size}")
    jl = np.asarray(ck["jloc"]).ravel()
    bad = np.flatnonzero((jl < 0) | (jl > nchunk - 2))
    if bad.size:
        raise RuntimeError(
            f"checkpoint: local well index {int(jl[bad[0]])} of particle {int(bad[0])} is outside the window "
            f"[0, {nchunk - 2}]"
        )
    if not np.all(np.isfinite(np.asarray(ck["y"]))):
        raise RuntimeError("checkpoint: yield positions are not finite")


def save_events(path, q, system=None):
    """Per-event records of `quasistatic` in the examples' output schema
    (examples/Line1d_Cuspy_Laplace.py:63-67: "x_frame", "f_frame", "S"; plus
    "inc"); `system` picks one row of a batch's records. A path ending in
    .h5 / .hdf5 is written as HDF5 (the examples' format, readable by their
    h5py scripts), any other as .npz."""
    pick = (lambda x: x) if system is None else (lambda x: np.asarray(x)[system])
    rec = dict(x_frame=pick(q["u_frame"]), f_frame=pick(q["f_frame"]), S=pick(q["S"]), inc=pick(q["inc"]))
    if str(path).endswith((".h5", ".hdf5")):
        from . import h5events

        h5events.write(path, rec)
    else:
        np.savez(path, **rec)


def load_events(path):
    """The records written by save_events ({"x_frame", "f_frame", "S", "inc"})."""
    if str(path).endswith((".h5", ".hdf5")):
        from . import h5events

        return h5events.read(path)
    with np.load(path) as f:
        return {k: f[k] for k in f.files}


def _check(rc):
    if rc != 0:
        raise RuntimeError(capi().fqsb_last_error().decode())


def make_desc(
    *,
    shape,
    seed,
    distribution,
    parameters,
    potential="Cuspy",
    interactions="Laplace1d",
    m=1.0,
    eta=0.0,
    mu=1.0,
    kappa=0.0,
    k_interactions=0.0,
    k_frame=0.0,
    dt=0.0,
    a1=0.0,
    a2=0.0,
    k2=0.0,
    k4=0.0,
    alpha=0.0,
    offset=-100.0,
    nchunk=5000,
    external=None,
    minimisation="Inertial",
    mean=0.0,
    stddev=0.0,
    seed_forcing=0,
    dinc_init=None,
    dinc=None,
):
    """Build a ``fqsb_desc`` from the reference's constructor arguments.

    ``external="RandomNormalForcing"`` adds the random forcing of
    Line1d.h:283-318 (``mean``, ``stddev``, ``seed_forcing``, ``dinc_init``,
    ``dinc``; the two arrays have the system's size). The description keeps
    references to the arrays it points to.
    """
    if distribution not in DISTRIBUTION:
        raise RuntimeError("Unknown distribution: " + str(distribution))
    shape = [int(s) for s in shape]
    d = Desc()
    d.rank = len(shape)
    d.potential = POTENTIAL[potential]
    d.interactions = INTERACTIONS[interactions]
    d.distribution = DISTRIBUTION[distribution]
    d.rows = shape[0]
    d.cols = shape[1] if len(shape) == 2 else 1
    d.seed = int(seed)
    d.nchunk = int(nchunk)
    d.m, d.eta, d.mu, d.kappa = float(m), float(eta), float(mu), float(kappa)
    d.k_frame, d.dt, d.k_interactions = float(k_frame), float(dt), float(k_interactions)
    d.a1, d.a2, d.k2, d.k4, d.alpha = float(a1), float(a2), float(k2), float(k4), float(alpha)
    d.offset = float(offset)
    par = list(parameters)
    if len(par) > 4:
        raise RuntimeError("Too many parameters for distribution")
    for i, p in enumerate(par):
        d.parameters[i] = float(p)
    d.nparameters = len(par)
    d.align_buffer = 2
    d.align_margin = 30
    if minimisation not in MINIMISATION:
        raise RuntimeError("unknown minimisation: " + str(minimisation))
    d.minimisation = MINIMISATION[minimisation]
    if external not in EXTERNAL:
        raise RuntimeError("unknown external forcing: " + str(external))
    d.external = EXTERNAL[external]
    if d.external:
        n = int(np.prod(shape))
        di = np.ascontiguousarray(np.broadcast_to(np.asarray(dinc_init, dtype=np.int64), [n]))
        dd = np.ascontiguousarray(np.broadcast_to(np.asarray(dinc, dtype=np.int64), [n]))
        d._keep = (di, dd)
        d.mean, d.stddev, d.seed_forcing = float(mean), float(stddev), int(seed_forcing)
        d.dinc_init = di.ctypes.data_as(_ip)
        d.dinc = dd.ctypes.data_as(_ip)
    return d


class Batch:
    """R replicas of one Line1d/Line2d system advanced together on the GPU."""

    def __init__(self, replicas, seed_stride=1, **kw):
        self._desc = make_desc(**kw)
        self._shape = [int(s) for s in kw["shape"]]
        self._N = int(np.prod(self._shape))
        self._R = int(replicas)
        h = C.c_void_p()
        _check(capi().fqsb_batch_create(C.byref(self._desc), self._R, int(seed_stride), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.fqsb_batch_destroy(h)
            self._h = None

    @property
    def replicas(self):
        return self._R

    @property
    def size(self):
        return self._N

    @property
    def shape(self):
        return list(self._shape)

    @property
    def nchunk(self):
        """Yield positions held per particle (the window of chunk.data)."""
        return int(self._desc.nchunk)

    @property
    def stream(self):
        """The hipStream_t (as an integer) the batch launches on."""
        return capi().fqsb_batch_stream(self._h)

    @property
    def longrange_method(self):
        """How LongRange1d forces are evaluated: "exact" (the reference's
        sequential sum) or "matrix" (circulant product on the matrix cores,
        the default for Cuspy batches)."""
        m = C.c_int()
        _check(capi().fqsb_batch_get_longrange_method(self._h, C.byref(m)))
        return {v: k for k, v in LONGRANGE_METHOD.items()}[m.value]

    @longrange_method.setter
    def longrange_method(self, name):
        if name not in LONGRANGE_METHOD:
            raise RuntimeError("unknown long-range method: " + str(name))
        _check(capi().fqsb_batch_set_longrange_method(self._h, LONGRANGE_METHOD[name]))

    @property
    def longrange_product_ms(self):
        """Average duration (ms) of one long-range product launch in the last
        timed timeSteps / flowSteps call of the matrix form."""
        ms = C.c_float()
        _check(capi().fqsb_batch_longrange_product_ms(self._h, C.byref(ms)))
        return ms.value

    def set_launch(self, ppt=0, threads=0):
        _check(capi().fqsb_batch_set_launch(self._h, int(ppt), int(threads)))

    @property
    def launch(self):
        p, t = C.c_int(), C.c_int()
        _check(capi().fqsb_batch_get_launch(self._h, C.byref(p), C.byref(t)))
        return p.value, t.value

    # launch shapes (particles per thread, threads per workgroup) the autotuner tries
    CANDIDATES = [(4, 1024), (4, 512), (8, 512), (8, 256), (16, 256), (16, 512), (8, 1024)]

    def autotune(self, v_frame=0.0, steps=500, candidates=None, repeats=2):
        """Time resident blocks per launch shape and keep the fastest.

        One untimed block first (clock ramp), then `repeats` timed blocks per
        candidate, best time kept. The trial steps advance the batch like
        flowSteps(steps, v_frame); use it during warm-up. Returns
        {(ppt, threads): ms}.
        """
        timings = {}
        self.flowSteps(steps, v_frame, block=steps)
        for ppt, threads in candidates or self.CANDIDATES:
            need = -(-self._N // ppt)  # threads that own particles
            if need > threads or need <= threads // 2:
                continue
            try:
                self.set_launch(ppt, threads)
            except RuntimeError:
                continue
            timings[(ppt, threads)] = min(
                self.flowSteps(steps, v_frame, block=steps, timed=True) for _ in range(repeats)
            )
        if not timings:
            self.set_launch(0, 0)
            return timings
        best = min(timings, key=timings.get)
        self.set_launch(*best)
        return timings

    def flowSteps(self, n, v_frame, block=0, timed=False):
        ms = C.c_float(0.0)
        _check(
            capi().fqsb_batch_flow_steps(
                self._h, int(n), float(v_frame), int(block), C.byref(ms) if timed else None
            )
        )
        return ms.value if timed else None

    def timeSteps(self, n, block=0, timed=False):
        ms = C.c_float(0.0)
        _check(capi().fqsb_batch_time_steps(self._h, int(n), int(block), C.byref(ms) if timed else None))
        return ms.value if timed else None

    def minimise(self, tol=1e-5, niter_tol=10, max_iter=int(1e9), max_iter_is_error=True):
        ret = np.zeros(self._R, dtype=np.uint64)
        _check(
            capi().fqsb_batch_minimise(
                self._h, tol, niter_tol, int(max_iter), int(max_iter_is_error), ret.ctypes.data_as(_up)
            )
        )
        return ret

    def minimise_truncate(self, i_n, A_truncate=0, S_truncate=0, tol=1e-5, niter_tol=10, max_iter=int(1e9),
                          max_iter_is_error=True):
        """minimise_truncate (detail.h:1833-1893) of every replica; i_n [R, N]."""
        i_n = np.ascontiguousarray(i_n, dtype=np.int64).reshape(self._R, self._N)
        ret = np.zeros(self._R, dtype=np.uint64)
        _check(
            capi().fqsb_batch_minimise_truncate(
                self._h, i_n.ctypes.data_as(_ip), int(A_truncate), int(S_truncate), float(tol), int(niter_tol),
                int(max_iter), int(bool(max_iter_is_error)), ret.ctypes.data_as(_up),
            )
        )
        return ret

    def trigger(self, p, eps, direction=1):
        """trigger (detail.h:1972-1977) of particle p[r] in replica r."""
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, dtype=np.uint64), [self._R]))
        _check(capi().fqsb_batch_trigger(self._h, p.ctypes.data_as(_up), float(eps), int(direction)))

    def advanceToFixedForce(self, f_frame, allow_plastic=False):
        """advanceToFixedForce (detail.h:1988-1995) to mean frame force f_frame[r]."""
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(f_frame, dtype=float), [self._R]))
        _check(capi().fqsb_batch_advance_to_fixed_force(self._h, f.ctypes.data_as(_dp), int(bool(allow_plastic))))

    def eventDrivenStep(self, eps, kick, direction=1):
        ret = np.zeros(self._R)
        _check(
            capi().fqsb_batch_event_driven_step(
                self._h, float(eps), int(bool(kick)), int(direction), ret.ctypes.data_as(_dp)
            )
        )
        return ret

    def quasistatic(self, eps, nevents, first_event=0, direction=1, kicks_only=False, tol=1e-5, niter_tol=10,
                    max_iter=int(1e9)):
        """The examples' quasistatic loop (examples/Line1d_Cuspy_Laplace.py:39-61)
        for every replica, each at its own pace on the device: events
        first_event .. first_event + nevents - 1 (0: u_frame = 0 and minimise;
        odd: eventDrivenStep(eps, False); even: eventDrivenStep(eps, True) and
        minimise; kicks_only: every event > 0 a kick and minimise). Returns a dict of [R, nevents] arrays: "S" (sum of the
        changes of chunk.index_at_align), "u_frame", "f_frame" (mean frame
        force) and "inc" after each event."""
        return _quasistatic(
            lambda *a: capi().fqsb_batch_quasistatic(self._h, *a), self._R, eps, nevents, first_event, direction,
            kicks_only, tol, niter_tol, max_iter,
        )

    @property
    def quasistatic_stats(self):
        """(ms, launches) of the last quasistatic(): time between HIP events
        around its device work and the number of driver launches."""
        ms = C.c_float()
        n = C.c_int64()
        _check(capi().fqsb_batch_quasistatic_stats(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # checkpoint / restart ---------------------------------------------------

    def _get_state(self, name):
        what, dtype = STATE[name]
        n, b = C.c_uint64(), C.c_int()
        _check(capi().fqsb_batch_state_size(self._h, what, C.byref(n), C.byref(b)))
        out = np.empty(n.value, dtype=dtype)
        assert out.itemsize == b.value
        _check(capi().fqsb_batch_get_state(self._h, what, out.ctypes.data_as(C.c_void_p)))
        return out

    def _set_state(self, name, value):
        what, dtype = STATE[name]
        n, b = C.c_uint64(), C.c_int()
        _check(capi().fqsb_batch_state_size(self._h, what, C.byref(n), C.byref(b)))
        v = np.ascontiguousarray(value, dtype=dtype).ravel()
        if v.size != n.value:
            raise RuntimeError(f"checkpoint item {name} has {v.size} values, the batch needs {n.value}")
        _check(capi().fqsb_batch_set_state(self._h, what, v.ctypes.data_as(C.c_void_p)))

    def checkpoint(self):
        """Everything that fixes the continuation bit for bit: u, v, a, the
        landscape windows with their pcg32 states, inc, u_frame, the activity
        counters and, with random forcing, its draws and state (the batched
        form of the reference's restart, test_Line1d.py:309-341). A dict of
        numpy arrays; `save` writes it as an .npz file."""
        ck = {k: self._get(k) for k in ("u", "v", "a")}
        items = ["y", "start", "jloc", "st_begin", "st_end", "inc", "u_frame", "qs_first", "qs_last"]
        if self._desc.external:
            items += ["f_thermal", "next", "rng_state"]
        for k in items:
            ck[k] = self._get_state(k)
        return ck

    def restore(self, ck):
        """Continue from a checkpoint of a batch of the same description. The
        whole checkpoint is checked first (validate_checkpoint), so a bad one
        raises without changing the batch."""
        validate_checkpoint(ck, self._R, self._shape, self.nchunk, bool(self._desc.external))
        for k in ("y", "start", "jloc", "st_begin", "st_end", "inc", "u_frame", "qs_first", "qs_last"):
            self._set_state(k, ck[k])
        if self._desc.external:
            for k in ("f_thermal", "next", "rng_state"):
                self._set_state(k, ck[k])
        for k in ("u", "v", "a"):
            self._set(k, ck[k])
        _check(capi().fqsb_batch_refresh(self._h))

    def save(self, path):
        np.savez(path, **self.checkpoint())

    def load(self, path):
        with np.load(path) as f:
            self.restore({k: f[k] for k in f.files})

    def _get(self, name):
        out = np.empty(self._R * self._N)
        _check(capi().fqsb_batch_get(self._h, FIELD[name], out.ctypes.data_as(_dp)))
        return out.reshape([self._R] + self._shape)

    def _set(self, name, value):
        v = np.ascontiguousarray(np.broadcast_to(value, [self._R] + self._shape), dtype=float).ravel()
        _check(capi().fqsb_batch_set(self._h, FIELD[name], v.ctypes.data_as(_dp)))

    u = property(lambda self: self._get("u"), lambda self, v: self._set("u", v))
    v = property(lambda self: self._get("v"), lambda self, v: self._set("v", v))
    a = property(lambda self: self._get("a"), lambda self, v: self._set("a", v))
    f = property(lambda self: self._get("f"))
    f_potential = property(lambda self: self._get("f_potential"))
    f_interactions = property(lambda self: self._get("f_interactions"))
    f_frame = property(lambda self: self._get("f_frame"))
    f_damping = property(lambda self: self._get("f_damping"))
    f_thermal = property(lambda self: self._get("f_thermal"))

    @property
    def u_frame(self):
        out = np.empty(self._R)
        _check(capi().fqsb_batch_get_u_frame(self._h, out.ctypes.data_as(_dp)))
        return out

    @property
    def inc(self):
        out = np.empty(self._R, dtype=np.int64)
        _check(capi().fqsb_batch_get_inc(self._h, out.ctypes.data_as(_ip)))
        return out

    @property
    def index_at_align(self):
        out = np.empty(self._R * self._N, dtype=np.int64)
        _check(capi().fqsb_batch_chunk_index_at_align(self._h, out.ctypes.data_as(_ip)))
        return out.reshape([self._R] + self._shape)
