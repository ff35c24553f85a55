"""Line-tile autotuner: the chunk length (points per thread) of each
direction's line plans, which fixes the chunks per line and with them the
line kernel and its block shape.  Solvers take it from a process-wide table
(filled by tune() or load()) and from a thread-count rule for shapes the
table does not hold; nothing is tuned implicitly, so a table always gives the
same plans and results.  No reference counterpart: the reference solves its
lines serially per rank (src/tridiag.jl:31-57, src/banded.jl:38-80)."""
import ctypes as C

from . import _lib as L


def tune(ctx, spec, reps: int = 5):
    """Times every candidate chunk length on spec's grid and stores the
    fastest per direction; returns (chunks, ms per repetition), 0 for
    inactive directions.  Solvers created afterwards use the choices."""
    chunks = (C.c_int * 3)()
    ms = (C.c_double * 3)()
    L.check(ctx.lib.cles_tune_solver(ctx.handle, C.byref(spec.desc), int(reps), chunks, ms))
    return tuple(chunks), tuple(ms)


def lookup(spec):
    """(chunks, tuned): the lengths a solver created from spec would use, and
    whether each comes from the table."""
    chunks = (C.c_int * 3)()
    tuned = (C.c_int * 3)()
    L.check(L.load().cles_tune_lookup(C.byref(spec.desc), chunks, tuned))
    return tuple(chunks), tuple(bool(t) for t in tuned)


def pin(spec, chunks):
    """Sets the chunk length of each direction for spec's grid (0: back to
    the rule, negative: unchanged)."""
    L.check(L.load().cles_tune_set(C.byref(spec.desc), (C.c_int * 3)(*[int(c) for c in chunks])))


def unpin(spec):
    pin(spec, (0, 0, 0))


def clear():
    L.check(L.load().cles_tune_clear())


def save(path: str):
    L.check(L.load().cles_tune_save(str(path).encode()))


def load(path: str):
    L.check(L.load().cles_tune_load(str(path).encode()))
