"""Ragged sets: systems of any classes, sizes and parameters advanced together.

No reference counterpart (the reference advances one system per object);
every operation applies to each system exactly as the single-system call of
the reference would (System::timeSteps, minimise, eventDrivenStep,
detail.h:1577-1960), so the results are bit-identical to separate systems.
Systems that share a kernel class and a size class run in one launch, and the
launches of different classes run concurrently.

Thin ctypes layer over the ``fqsb_ragged_*`` functions of ``libfqsb.so``
(include/fqsb.h).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .batch import FIELD, _check, _dp, _ip, _quasistatic, _up, capi, make_desc

_u8p = C.POINTER(C.c_uint8)

_CLASSES = {
    # class name (Line1d / Line2d of the reference) -> make_desc arguments
    "System_Cuspy_Laplace": dict(potential="Cuspy", interactions="Laplace1d"),
    "System_Cuspy_Laplace_Nopassing": dict(potential="Cuspy", interactions="Laplace1d", minimisation="Overdamped"),
    "System_SemiSmooth_Laplace": dict(potential="SemiSmooth", interactions="Laplace1d"),
    "System_Smooth_Laplace": dict(potential="Smooth", interactions="Laplace1d"),
    "System_Cuspy_Quartic": dict(potential="Cuspy", interactions="Quartic1d"),
    "System_SemiSmooth_Quartic": dict(potential="SemiSmooth", interactions="Quartic1d"),
    "System_Smooth_Quartic": dict(potential="Smooth", interactions="Quartic1d"),
    "System_Cuspy_QuarticGradient": dict(potential="Cuspy", interactions="QuarticGradient1d"),
    "System_Cuspy_LongRange": dict(potential="Cuspy", interactions="LongRange1d"),
    "Particles.System_Cuspy": dict(potential="Cuspy", interactions=None),
    "Particles.System_SemiSmooth": dict(potential="SemiSmooth", interactions=None),
    "Particles.System_Smooth": dict(potential="Smooth", interactions=None),
    "Line2d.System_Cuspy_Laplace": dict(potential="Cuspy", interactions="Laplace2d"),
    "Line2d.System_Cuspy_QuarticGradient": dict(potential="Cuspy", interactions="QuarticGradient2d"),
}


class Ragged:
    """A set of systems, each given as ``(class_name, kwargs)`` with the
    reference's constructor arguments (``Line1d.System_*``, with or without the
    ``Line1d.`` prefix; ``Line2d.`` prefix for the 2-d classes), e.g. ``("System_Smooth_Laplace", dict(m=1, ...))``."""

    def __init__(self, systems):
        self._descs = []
        self._shapes = []
        for name, kw in systems:
            key = name[len("Line1d."):] if str(name).startswith("Line1d.") else name
            if key not in _CLASSES:
                raise RuntimeError("unknown system class: " + str(name))
            self._descs.append(make_desc(**_CLASSES[key], **kw))
            self._shapes.append([int(s) for s in kw["shape"]])
        n = len(self._descs)
        arr = (type(self._descs[0]) * n)(*self._descs)
        self._arr = arr
        h = C.c_void_p()
        _check(capi().fqsb_ragged_create(arr, n, C.byref(h)))
        self._h = h
        self._systems = [(name, dict(kw)) for name, kw in systems]
        self._serial = False
        self._n = n
        self._sizes = [int(np.prod(s)) for s in self._shapes]
        self._offsets = np.concatenate([[0], np.cumsum(self._sizes)]).astype(int)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None:
            try:
                capi().fqsb_ragged_destroy(h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return self._n

    @property
    def sizes(self):
        return list(self._sizes)

    @property
    def groups(self):
        """Number of launch groups (kernel class x size class)."""
        return int(capi().fqsb_ragged_groups(self._h))

    def _mask(self, active):
        if active is None:
            return None
        m = np.ascontiguousarray(np.asarray(active, dtype=bool).astype(np.uint8))
        if m.size != self._n:
            raise RuntimeError("active must have one entry per system")
        return m

    def timeSteps(self, n):
        _check(capi().fqsb_ragged_time_steps(self._h, int(n)))

    def flowSteps(self, n, v_frame):
        _check(capi().fqsb_ragged_flow_steps(self._h, int(n), float(v_frame)))

    def minimise(self, tol=1e-5, niter_tol=10, max_iter=int(1e9), max_iter_is_error=True, active=None):
        ret = np.zeros(self._n, dtype=np.uint64)
        m = self._mask(active)
        _check(
            capi().fqsb_ragged_minimise(
                self._h, float(tol), int(niter_tol), int(max_iter), int(bool(max_iter_is_error)),
                None if m is None else m.ctypes.data_as(_u8p), ret.ctypes.data_as(_up),
            )
        )
        return ret

    def eventDrivenStep(self, eps, kick, direction=1, active=None):
        ret = np.zeros(self._n)
        m = self._mask(active)
        _check(
            capi().fqsb_ragged_event_driven_step(
                self._h, float(eps), int(bool(kick)), int(direction),
                None if m is None else m.ctypes.data_as(_u8p), ret.ctypes.data_as(_dp),
            )
        )
        return ret

    def quasistatic(self, eps, nevents, first_event=0, direction=1, kicks_only=False, tol=1e-5, niter_tol=10,
                    max_iter=int(1e9), active=None):
        """The examples' quasistatic loop for every active system, each at its
        own pace (batch.Batch.quasistatic); [systems, nevents] arrays."""
        return _quasistatic(
            lambda *a: capi().fqsb_ragged_quasistatic(self._h, *a), self._n, eps, nevents, first_event, direction,
            kicks_only, tol, niter_tol, max_iter, masked=True, mask=self._mask(active),
        )

    # launch groups and their tuning ---------------------------------------

    PATHS = {0: "resident", 1: "tiles", 2: "streaming", 3: "nopassing-resident", 4: "nopassing-tiles"}

    def group_info(self, g):
        """Launch group g: its systems (indices), largest size, path and
        current launch shape / block length."""
        info = np.empty(13, dtype=np.int64)
        _check(capi().fqsb_ragged_group_info(self._h, int(g), info.ctypes.data_as(_ip)))
        idx = np.empty(int(info[0]), dtype=np.uint64)
        _check(capi().fqsb_ragged_group_systems(self._h, int(g), idx.ctypes.data_as(_up)))
        ppt = (C.c_int * 32)()
        thr = (C.c_int * 32)()
        cnt = C.c_int()
        _check(capi().fqsb_ragged_group_shapes(self._h, int(g), ppt, thr, 32, C.byref(cnt)))
        return {
            "systems": idx.astype(int),
            "N": int(info[1]),
            "particles": int(info[11]),
            "path": self.PATHS.get(int(info[2]), str(info[2])),
            "launch": (int(info[3]), int(info[4])),
            "block": int(info[5]),
            "block_max": int(info[12]),
            "key": "-".join(str(int(x)) for x in info[6:11]) + f"-{int(np.ceil(np.log2(max(int(info[1]), 1))))}",
            "shapes": [(int(ppt[i]), int(thr[i])) for i in range(min(cnt.value, 32))],
        }

    def group_times(self):
        """[(ms, particle_steps)] per group since the last reset: the time of
        the group's part of every operation on its stream and the work it did."""
        out = []
        for g in range(self.groups):
            ms, ps = C.c_double(), C.c_double()
            _check(capi().fqsb_ragged_group_time(self._h, g, C.byref(ms), C.byref(ps)))
            out.append((ms.value, ps.value))
        return out

    def reset_group_times(self):
        _check(capi().fqsb_ragged_reset_group_time(self._h))

    @property
    def serial(self):
        """Run the launch groups one after the other (True) instead of
        concurrently on their streams (False, the default); results are the
        same, and each group's time is then its own."""
        return self._serial

    @serial.setter
    def serial(self, value):
        _check(capi().fqsb_ragged_set_serial(self._h, int(bool(value))))
        self._serial = bool(value)

    def set_group_launch(self, g, ppt, threads):
        _check(capi().fqsb_ragged_set_group_launch(self._h, int(g), int(ppt), int(threads)))

    def set_group_block(self, g, steps):
        _check(capi().fqsb_ragged_set_group_block(self._h, int(g), int(steps)))

    def autotune(self, eps=1e-3, events=2, blocks=(16, 32, 64, 128), cache=None, smooth=None, min_particles=0,
                 margin=0.1):
        """Choose each launch group's launch shape (resident groups) or block
        length (tiled groups) by timing a few events of the examples'
        quasistatic protocol per candidate, on a copy of this set built from
        the same descriptions (so this set's own state is left as it is);
        results do not depend on the choice. Candidates run successive events
        (of different avalanche sizes), so they are compared by the group's
        time per particle-step (events on its stream, the other groups idle).
        `smooth` marks the systems that take kicks only (their
        maxUniformDisplacement raises). With `cache` (a JSON file), choices
        found there for a group's key (class and size class) are applied
        without trials and new ones are added. Groups of fewer than
        `min_particles` particles keep their default. A trial runs the group
        alone, while a set's groups normally share the GPU (and its few
        hardware queues), where more and shorter launches cost more than the
        trial shows: a candidate replaces the group's default only if it is
        faster by more than `margin` (a fraction). Returns {group: {"key",
        "path", "choice", "trials_ms", "trials_ns_per_particle_step"}}."""
        import json
        import os
        import time

        known = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        smooth = np.zeros(self._n, dtype=bool) if smooth is None else np.asarray(smooth, dtype=bool)
        out = {}
        trial = None
        first = 1
        for g in range(self.groups):
            info = self.group_info(g)
            key = info["key"]
            if info["particles"] < min_particles:
                continue
            if info["path"] == "resident" and len(info["shapes"]) > 1:
                cands = [("launch", sh) for sh in info["shapes"]]
            elif info["path"] == "tiles" and info["block"] > 0:
                # block lengths within what the tiles' halos allow (lattices:
                # a few rows of halo, so shorter blocks as well)
                kmax = info["block_max"]
                ks = {min(k, kmax) for k in blocks}
                if kmax <= min(blocks):
                    ks |= {max(1, kmax // 4), max(1, kmax // 2)}
                cands = [("block", k) for k in sorted(ks)]
                if len(cands) < 2:
                    continue
            else:
                continue
            if key in known:
                kind, val = known[key]["choice"]
                self._apply(g, kind, val)
                out[g] = {"key": key, "path": info["path"], "choice": [kind, val], "trials_ms": None, "cached": True}
                continue
            if trial is None:
                trial = Ragged(self._systems)
                trial.quasistatic(eps, 1)  # event 0 of every system
            act = np.zeros(self._n, dtype=bool)
            act[info["systems"]] = True
            default = ("launch", tuple(info["launch"])) if info["path"] == "resident" else ("block", info["block"])
            if default not in cands:
                cands.append(default)
            times, rate = {}, {}
            for kind, val in cands:
                trial._apply(g, kind, val)
                trial.reset_group_times()
                t0 = time.perf_counter()
                if np.any(act & ~smooth):
                    trial.quasistatic(eps, 2 * events, first_event=first, active=act & ~smooth)
                if np.any(act & smooth):
                    trial.quasistatic(eps, events, first_event=(first + 1) // 2, kicks_only=True, active=act & smooth)
                name = f"{kind}:{val}"
                times[name] = (time.perf_counter() - t0) * 1e3
                ms, ps = trial.group_times()[g]
                rate[name] = ms * 1e6 / ps if ps > 0 else times[name] * 1e6
                first += 2 * events
            best = min(rate, key=rate.get)
            base = f"{default[0]}:{default[1]}"
            if rate[best] > (1.0 - margin) * rate[base]:
                best = base
            kind, val = best.split(":", 1)
            val = [int(x) for x in val.strip("()").split(",")] if kind == "launch" else int(val)
            self._apply(g, kind, val)
            known[key] = {"choice": [kind, val]}
            out[g] = {"key": key, "path": info["path"], "choice": [kind, val], "trials_ms": times,
                      "trials_ns_per_particle_step": rate, "cached": False}
        if cache:
            with open(cache, "w") as f:
                json.dump(known, f, indent=1, sort_keys=True)
        return out

    def _apply(self, g, kind, val):
        if kind == "launch":
            self.set_group_launch(g, *val)
        else:
            self.set_group_block(g, val)

    def _split(self, flat):
        return [flat[a:b].reshape(s) for a, b, s in zip(self._offsets[:-1], self._offsets[1:], self._shapes)]

    def _get(self, name):
        out = np.empty(self._offsets[-1])
        _check(capi().fqsb_ragged_get(self._h, FIELD[name], out.ctypes.data_as(_dp)))
        return self._split(out)

    def _set(self, name, values):
        flat = np.ascontiguousarray(np.concatenate([np.ravel(np.asarray(v, dtype=float)) for v in values]))
        if flat.size != self._offsets[-1]:
            raise RuntimeError("one array per system of the system's size is needed")
        _check(capi().fqsb_ragged_set(self._h, FIELD[name], flat.ctypes.data_as(_dp)))

    u = property(lambda self: self._get("u"), lambda self, v: self._set("u", v))
    v = property(lambda self: self._get("v"), lambda self, v: self._set("v", v))
    a = property(lambda self: self._get("a"), lambda self, v: self._set("a", v))
    f = property(lambda self: self._get("f"))
    f_potential = property(lambda self: self._get("f_potential"))
    f_interactions = property(lambda self: self._get("f_interactions"))
    f_frame = property(lambda self: self._get("f_frame"))
    f_damping = property(lambda self: self._get("f_damping"))

    @property
    def u_frame(self):
        out = np.empty(self._n)
        _check(capi().fqsb_ragged_get_u_frame(self._h, out.ctypes.data_as(_dp)))
        return out

    @property
    def inc(self):
        out = np.empty(self._n, dtype=np.int64)
        _check(capi().fqsb_ragged_get_inc(self._h, out.ctypes.data_as(_ip)))
        return out

    @property
    def last_steps(self):
        """Steps (or no-passing sweeps) each system took in the last run."""
        out = np.empty(self._n, dtype=np.uint64)
        _check(capi().fqsb_ragged_last_steps(self._h, out.ctypes.data_as(_up)))
        return out

    @property
    def index_at_align(self):
        out = np.empty(self._offsets[-1], dtype=np.int64)
        _check(capi().fqsb_ragged_chunk_index_at_align(self._h, out.ctypes.data_as(_ip)))
        return self._split(out)
