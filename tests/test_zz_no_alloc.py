"""cles_batch_run_noalloc, the batch run! that allocates no device memory,
against cles_batch_run on identical batches, called through the C ABI: two
calls of each in a row, the second with a smaller batch and a shorter dt
history, must give bitwise the same dt histories, step counts, times,
statuses and final states.  The second call would show records left over
from the first if the shared array were not cleared.  Also the refusal of a
dt history larger than the array."""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _run(ctx, specs, nrec, steps, noalloc):
    from compactles_hip import _lib as L
    b = cases.tube_batch(ctx, specs)
    fn = ctx.lib.cles_batch_run_noalloc if noalloc else ctx.lib.cles_batch_run
    dts = np.full((len(specs), nrec), np.nan)
    taken = C.c_int64()
    L.check(fn(b.handle, steps, nrec, dts.ctypes.data_as(L.DP), C.byref(taken)))
    t, st, status = b.state()
    Q = [b.download(i) for i in range(len(specs))]
    b.close()
    return dts, taken.value, t, st, status, Q


def test_noalloc_run_equals_allocating_run(ctx):
    big = cases.tube_specs(24, seed=7, nmin=64, nmax=400)
    small = cases.tube_specs(6, seed=8, nmin=64, nmax=300)
    for specs, nrec, steps in ((big, 30, 25), (small, 12, 10)):
        print(f"case noalloc: {len(specs)} tubes, nrec {nrec}, {steps} steps", flush=True)
        got = _run(ctx, specs, nrec, steps, True)
        ref = _run(ctx, specs, nrec, steps, False)
        assert np.array_equal(got[0], ref[0])            # dt histories, zero past the steps taken
        assert np.all(got[0][:, :steps] > 0) and np.all(got[0][:, steps:] == 0)
        assert got[1] == ref[1] == steps
        for a, b in zip(got[2:5], ref[2:5]):
            assert np.array_equal(a, b)
        for qa, qb in zip(got[5], ref[5]):
            assert np.array_equal(qa, qb)


def test_noalloc_refuses_a_history_beyond_its_array(ctx):
    from compactles_hip import _lib as L
    b = cases.tube_batch(ctx, cases.tube_specs(2, seed=9, nmin=64, nmax=128))
    nrec = (1 << 19) + 1                                  # 2 * nrec > CLES_NOALLOC_DTS_MAX = 2^20
    dts = np.zeros((2, nrec))
    taken = C.c_int64(-1)
    rc = ctx.lib.cles_batch_run_noalloc(b.handle, 1, nrec, dts.ctypes.data_as(L.DP), C.byref(taken))
    assert rc == 1 and b"CLES_NOALLOC_DTS_MAX" in ctx.lib.cles_last_error() and taken.value == -1
    t, st, status = b.state()
    assert np.all(st == 0)                                # nothing ran
    b.close()
