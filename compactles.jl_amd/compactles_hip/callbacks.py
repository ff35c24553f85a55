"""Step-boundary callbacks of run! (src/callbacks.jl): a Trigger decides when,
an effect what, both evaluated on the host between completed steps (the
library calls back through cles_run_hooks).

Triggers follow callbacks.jl: AtTime (callbacks.jl:58-98), EveryTime
(:100-158), EveryStep (:160-175), WhenState (:177-206); Callback pairs one
with effect(solver, Q) (:208-250), whose `True` return stops run!; a tuple or
list of callbacks runs all of them after each step; a bare callable runs
after every step with its return value ignored.  ProgressLog is
callbacks.jl:262-330.

Triggers see a RunClock (t, step, dt_prev, cfl of the run in progress);
effects and WhenState conditions receive the solver and Q, a lazy view of the
device state (np.asarray(Q) downloads it at that step).
"""
import ctypes as C
import math
import sys
import time

import numpy as np


def _land_tol(t):
    """callbacks.jl:98: 8 eps(max(|t|, 1))."""
    return 8.0 * float(np.spacing(max(abs(t), 1.0)))


class RunClock:
    """The run's clock as the triggers read it (solver.t, solver.step, ...)."""

    def __init__(self, t=0.0, step=0):
        self.t, self.step = float(t), int(step)
        self.dt_prev = 0.0
        self.wall_step = 0.0
        self.wall_total = 0.0


class Trigger:
    """callbacks.jl:27-56."""

    def next_time(self, clock):
        return math.inf

    def fired(self, clock, solver, Q):
        raise NotImplementedError

    def rewind(self, t, step):
        pass


class AtTime(Trigger):
    """Fires once at each listed time; run! lands a step on each (callbacks.jl:58-98)."""

    def __init__(self, times):
        if np.isscalar(times):
            times = [times]
        self.times = sorted(float(x) for x in times)
        self.next = 0

    def next_time(self, clock):
        return self.times[self.next] if self.next < len(self.times) else math.inf

    def fired(self, clock, solver, Q):
        target = self.next_time(clock)
        if not math.isfinite(target):
            return False
        if clock.t < target - _land_tol(target):
            return False
        self.next += 1
        return True

    def rewind(self, t, step):
        for i, target in enumerate(self.times):
            if target > t + _land_tol(target):
                self.next = i
                return
        self.next = len(self.times)


class EveryTime(Trigger):
    """Fires at start + n*interval (callbacks.jl:100-158)."""

    def __init__(self, interval, start=0.0):
        if not interval > 0:
            raise ValueError("EveryTime interval must be positive")
        self.interval, self.start = float(interval), float(start)
        self.next = math.nan

    def _instant_after(self, t):
        n = max(math.floor((t - self.start) / self.interval) + 1, 0.0)
        inst = self.start + n * self.interval
        while inst <= t + _land_tol(t):
            n += 1
            inst = self.start + n * self.interval
        return inst

    def next_time(self, clock):
        if math.isnan(self.next):
            self.next = self._instant_after(clock.t)
        return self.next

    def fired(self, clock, solver, Q):
        target = self.next_time(clock)
        if clock.t < target - _land_tol(target):
            return False
        self.next = self._instant_after(clock.t)
        return True

    def rewind(self, t, step):
        self.next = math.nan


class EveryStep(Trigger):
    """Fires every `interval` completed steps (callbacks.jl:160-175)."""

    def __init__(self, interval=1):
        if int(interval) < 1:
            raise ValueError("EveryStep interval must be >= 1")
        self.interval = int(interval)

    def fired(self, clock, solver, Q):
        return clock.step % self.interval == 0


class WhenState(Trigger):
    """Fires when condition(solver, Q) first holds, or whenever it holds with
    once=False (callbacks.jl:177-206; one rank, so no reduction)."""

    def __init__(self, condition, once=True):
        self.condition, self.once, self.done = condition, bool(once), False

    def fired(self, clock, solver, Q):
        if self.done:
            return False
        hit = bool(self.condition(solver, Q))
        if hit and self.once:
            self.done = True
        return hit


class Callback:
    """A trigger and effect(solver, Q); a True return stops run! (callbacks.jl:208-222)."""

    def __init__(self, trigger, effect):
        self.trigger, self.effect = trigger, effect


def callback_next_time(cb, clock):
    """callbacks.jl:215-222."""
    if cb is None:
        return math.inf
    if isinstance(cb, Callback):
        return cb.trigger.next_time(clock)
    if isinstance(cb, (tuple, list)):
        return min((callback_next_time(c, clock) for c in cb), default=math.inf)
    return math.inf


def rewind_callbacks(cb, t, step):
    """callbacks.jl:224-231."""
    if isinstance(cb, Callback):
        cb.trigger.rewind(t, step)
    elif isinstance(cb, (tuple, list)):
        for c in cb:
            rewind_callbacks(c, t, step)


def run_callbacks(cb, clock, solver, Q):
    """callbacks.jl:233-254: every callback runs (no short circuit); True stops."""
    if cb is None:
        return False
    if isinstance(cb, Callback):
        if not cb.trigger.fired(clock, solver, Q):
            return False
        return cb.effect(solver, Q) is True
    if isinstance(cb, (tuple, list)):
        stop = False
        for c in cb:
            stop = run_callbacks(c, clock, solver, Q) or stop
        return stop
    cb(solver, Q)   # a bare callable: every step, return value ignored
    return False


class ProgressEffect:
    def __init__(self, quantity, label, io, tfinal):
        self.quantity, self.label, self.io, self.tfinal = quantity, label, io, float(tfinal)

    def __call__(self, solver, Q):
        clock = solver.clock
        value = None if self.quantity is None else self.quantity(solver, Q)
        line = (f"step {clock.step:7d}  t = {clock.t:11.5e}  dt = {clock.dt_prev:10.4e}  "
                f"{clock.wall_step:8.4g} s/step")
        if math.isfinite(self.tfinal) and self.tfinal > 0:
            frac = min(max(clock.t / self.tfinal, 0.0), 1.0)
            mean = clock.wall_total / clock.step if clock.step > 0 else 0.0
            rem = mean * clock.step * (1 - frac) / frac if frac > 0 else math.nan
            line += f"  {100 * frac:5.1f}%  eta {_duration(rem)}"
        if value is not None:
            line += f"  {value:.8e}" if not self.label else f"  {self.label} = {value:.8e}"
        print(line, file=self.io, flush=True)
        return False


def _duration(seconds):
    if not math.isfinite(seconds):
        return "  --  "
    if seconds < 90:
        return f"{seconds:4.0f}s"
    if seconds < 5400:
        return f"{seconds / 60:4.1f}m"
    return f"{seconds / 3600:4.1f}h"


def ProgressLog(every=10, quantity=None, label="", tfinal=math.inf, io=None):
    """callbacks.jl:262-330 (one rank: no imbalance columns)."""
    return Callback(EveryStep(every), ProgressEffect(quantity, str(label), io or sys.stdout, tfinal))


# ---------------------------------------------------------------------------
# The C hook table (cles_run_hooks; the oracle's or_run_hooks has the same layout)
# ---------------------------------------------------------------------------
NEXT_TIME = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_int64)
AFTER_STEP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_int64, C.c_double)
REWIND = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_int64)


class RunHooks(C.Structure):
    _fields_ = [("user", C.c_void_p), ("next_time", NEXT_TIME), ("after_step", AFTER_STEP), ("rewind", REWIND)]


class LazyState:
    """Q for effects: np.asarray(Q) downloads the state at that step."""

    def __init__(self, fetch):
        self._fetch = fetch

    def __array__(self, dtype=None, copy=None):
        a = self._fetch()
        return a if dtype is None else a.astype(dtype)


class HookAdapter:
    """Turns a callback (None, Callback, tuple/list, bare callable) into a
    RunHooks table.  An exception in a trigger or effect stops the run and is
    re-raised by check() after it returns."""

    def __init__(self, callback, solver, Q, t0, step0):
        self.cb, self.solver, self.Q = callback, solver, Q
        self.clock = RunClock(t0, step0)
        self.rewinds = []
        self.error = None
        self._wall = time.perf_counter()
        solver.clock = self.clock
        self.hooks = RunHooks(None, NEXT_TIME(self._next_time), AFTER_STEP(self._after_step), REWIND(self._rewind))

    def _next_time(self, user, t, step):
        if self.error is not None:
            return math.inf
        try:
            self.clock.t, self.clock.step = t, step
            return float(callback_next_time(self.cb, self.clock))
        except BaseException as e:   # noqa: BLE001 - re-raised by check()
            self.error = e
            return math.inf

    def _after_step(self, user, t, step, dt):
        if self.error is not None:
            return 1
        try:
            now = time.perf_counter()
            c = self.clock
            c.t, c.step, c.dt_prev = t, step, dt
            c.wall_step = now - self._wall
            c.wall_total += c.wall_step
            self._wall = now
            stop = run_callbacks(self.cb, c, self.solver, self.Q)
            self._wall = time.perf_counter()   # callbacks are outside the step time (timestep.jl:291-296)
            return 1 if stop else 0
        except BaseException as e:   # noqa: BLE001
            self.error = e
            return 1

    def _rewind(self, user, t, step):
        try:
            self.rewinds.append((t, step))
            self.clock.t, self.clock.step = t, step
            rewind_callbacks(self.cb, t, step)
        except BaseException as e:   # noqa: BLE001
            self.error = e

    def check(self):
        if self.error is not None:
            raise self.error
