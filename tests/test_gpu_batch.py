"""GPU parity of batched one-dimensional instances (many shock tubes of
different lengths in one set of launches) against the oracle run on each tube
alone, through the C ABI; plus the batch's own contracts: an instance's
result is the same alone and in any mix, each stops at its own tfinal, and a
failing instance is frozen with its status while the others complete.
Bar: FP64 dQ <= 1e-12; after a run, dt history <= 1e-12, state <= 1e-10."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _state_err(got, ref):
    """max |got - ref| / scale per component; momentum shares one scale."""
    C = len(ref)
    Ns = C - 4
    scale = [np.abs(r).max() for r in ref]
    ms = max(scale[Ns:Ns + 3])
    for c in range(Ns, Ns + 3):
        scale[c] = ms
    return max(np.abs(got[c] - ref[c]).max() / max(scale[c], 1e-300) for c in range(C))


SPECS = cases.tube_specs(5, seed=7, nmin=128, nmax=1100) + [{"n": 2048, "pL": 9.0, "rhoL": 3.0, "pR": 1.0,
                                                              "rhoR": 1.0}]


def test_batch_rhs_parity(ctx):
    b = cases.tube_batch(ctx, SPECS)
    b.compute_rhs(False)
    for i, sp in enumerate(SPECS):
        o = cases.tube_oracle(sp)
        o.compute_rhs(False)
        ref = [o.interior(o.dQ[..., c])[:, 0, 0] for c in range(o.n_cons)]
        got = [b.field(i, f"dQ[{c + 1}]") for c in range(o.n_cons)]
        err = _state_err(got, ref)
        assert err <= 1e-12, (i, sp["n"], err)


def test_batch_run_parity(ctx):
    nsteps = 25
    b = cases.tube_batch(ctx, SPECS, nmax=nsteps)
    taken, dts = b.run(record=nsteps)
    t, steps, status = b.state()
    assert (status == 0).all() and (steps == nsteps).all()
    ke = b.kinetic_energy()
    for i, sp in enumerate(SPECS):
        o = cases.tube_oracle(sp)
        d1 = o.run(1e9, nsteps, record=nsteps)
        assert np.abs(dts[i] - d1).max() <= 1e-12 * d1.max(), (i, sp["n"])
        assert abs(t[i] - o.time[0]) <= 1e-12 * o.time[0]
        ref = [o.interior(o.Q[..., c])[:, 0, 0] for c in range(o.n_cons)]
        got = [b.field(i, f"Q[{c + 1}]") for c in range(o.n_cons)]
        assert _state_err(got, ref) <= 1e-10, (i, sp["n"], _state_err(got, ref))
        ke_ref = o.diagnostics()[0]
        assert abs(ke[i] - ke_ref) <= 1e-10 * ke_ref, (i, ke[i], ke_ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_step_graph_equals_host(ctx, precision):
    """Iterations replayed from the captured 8-iteration graph against the
    same batch launched from the host: dt histories and states bitwise equal,
    over a first call with 2 replays and 5 host-issued iterations, a second
    call that replays the same graph, and a third with another history length
    (a new capture); one tube stops on its tfinal inside a replay."""
    specs = cases.tube_specs(7, seed=5, nmin=128, nmax=2048)
    tf = [1e9] * len(specs)
    tf[2] = 0.003
    a = cases.tube_batch(ctx, specs, precision=precision, tfinal=tf)
    b = cases.tube_batch(ctx, specs, precision=precision, tfinal=tf)
    a.set_run_mode("step_graph")
    b.set_run_mode("host_dt")
    for steps, rec in ((21, 21), (16, 16), (9, 4)):
        na, da = a.run(steps, record=rec)
        nb, db = b.run(steps, record=rec)
        assert na == nb and np.array_equal(da, db), (steps, na, nb)
        ta, sa, xa = a.state()
        tb, sb, xb = b.state()
        assert np.array_equal(ta, tb) and np.array_equal(sa, sb) and np.array_equal(xa, xb)
        for i in range(len(specs)):
            for c in range(a.n_cons):
                assert np.array_equal(a.field(i, f"Q[{c + 1}]"), b.field(i, f"Q[{c + 1}]")), (steps, i, c)
    assert sa[2] < sa[0]   # the short tube stopped on its tfinal
    with pytest.raises(ValueError):
        a.set_run_mode("graph")
    a.close()
    b.close()


def test_batch_same_alone_and_in_mix(ctx):
    specs = cases.tube_specs(12, seed=11, nmin=128, nmax=2048)
    nsteps = 15
    full = cases.tube_batch(ctx, specs, nmax=nsteps)
    full.run()
    j = 5
    alone = cases.tube_batch(ctx, [specs[j]], nmax=nsteps)
    alone.run()
    order = list(reversed(range(len(specs))))
    mixed = cases.tube_batch(ctx, [specs[k] for k in order], nmax=nsteps)
    mixed.run()
    jm = order.index(j)
    for c in range(6):
        a = alone.field(0, f"Q[{c + 1}]")
        assert np.array_equal(a, full.field(j, f"Q[{c + 1}]")), c
        assert np.array_equal(a, mixed.field(jm, f"Q[{c + 1}]")), c


def test_batch_own_tfinal(ctx):
    specs = cases.tube_specs(4, seed=3, nmin=200, nmax=600)
    tf = [2e-3, 5e-3, 1e-3, 4e-3]
    b = cases.tube_batch(ctx, specs, tfinal=tf)
    b.run()
    t, steps, status = b.state()
    assert (status == 0).all()
    assert np.allclose(t, tf, rtol=1e-12, atol=0)
    for i, sp in enumerate(specs):
        o = cases.tube_oracle(sp)
        o.run(tf[i], 10 ** 6)
        assert steps[i] == o.time[1]
        ref = [o.interior(o.Q[..., c])[:, 0, 0] for c in range(6)]
        got = [b.field(i, f"Q[{c + 1}]") for c in range(6)]
        assert _state_err(got, ref) <= 1e-10


def test_batch_failure_is_isolated(ctx):
    """An instance whose state has negative density stops with the
    negative-density status at its first step; the others are unaffected."""
    specs = cases.tube_specs(3, seed=5, nmin=150, nmax=400)
    nsteps = 10
    b = cases.tube_batch(ctx, specs, nmax=nsteps)
    bad = b.download(1)
    bad[4 + 10:4 + 20, 0, 0, 0] = -1.0          # rho Y_1 < 0 over a few nodes
    b.upload(1, bad)
    b.run()
    t, steps, status = b.state()
    assert status[1] == 5 and steps[1] == 0 and t[1] == 0.0
    assert status[0] == 0 and status[2] == 0 and steps[0] == nsteps and steps[2] == nsteps
    ref = cases.tube_batch(ctx, [specs[0], specs[2]], nmax=nsteps)
    ref.run()
    for k, i in ((0, 0), (1, 2)):
        for c in range(6):
            assert np.array_equal(ref.field(k, f"Q[{c + 1}]"), b.field(i, f"Q[{c + 1}]"))
    # the failed instance keeps its state (its ends re-imposed by the Dirichlet condition)
    after = b.download(1)
    assert np.array_equal(after[4 + 1:4 + specs[1]["n"] - 1], bad[4 + 1:4 + specs[1]["n"] - 1])


def test_batch_fp32_mixed_error(ctx):
    nsteps = 25
    b = cases.tube_batch(ctx, SPECS, nmax=nsteps, precision="fp32")
    b.run()
    ke = b.kinetic_energy()
    errs = []
    for i, sp in enumerate(SPECS):
        o = cases.tube_oracle(sp)
        o.run(1e9, nsteps)
        ref = [o.interior(o.Q[..., c])[:, 0, 0] for c in range(o.n_cons)]
        got = [b.field(i, f"Q[{c + 1}]") for c in range(o.n_cons)]
        ke_ref = o.diagnostics()[0]
        errs.append((abs(ke[i] - ke_ref) / ke_ref, _state_err(got, ref)))
    e_ke = max(e[0] for e in errs)
    e_q = max(e[1] for e in errs)
    print(f"\nFP32-mixed batched tubes, {nsteps} steps vs FP64 oracle: KE {e_ke:.2e}  Q {e_q:.2e}")
    assert e_ke <= 1e-5 and e_q <= 1e-4


def test_batch_upload_refuses_a_transverse_layout(ctx):
    """An instance is held as (n + 2H, 1, 1) with halo (H, 0, 0): a transverse
    extent or halo would address another row, so it is refused."""
    import compactles_hip._lib as L
    b = cases.tube_batch(ctx, cases.tube_specs(2, seed=3, nmin=64, nmax=96))
    n = b.instances[0].n
    for padded, halo in (((n + 8, 2, 1), (4, 0, 0)), ((n + 8, 1, 1), (4, 1, 0)), ((n + 8, 1, 3), (4, 0, 1))):
        Q = np.zeros(padded + (b.n_cons,), order="F")
        rc = ctx.lib.cles_batch_upload(b.handle, 0, Q.ctypes.data_as(L.DP), L.I64_3(*padded), L.INT_3(*halo))
        assert rc == L.ERR_INVALID
        assert "laid out" in ctx.lib.cles_last_error().decode()
        rc = ctx.lib.cles_batch_download(b.handle, 0, Q.ctypes.data_as(L.DP), L.I64_3(*padded), L.INT_3(*halo))
        assert rc != 0
    b.close()
