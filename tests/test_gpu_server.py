"""The serving loop on the GPU: a Poisson-like mix of Taylor-Green and 2-D
shock / interface requests admitted by footprint and run concurrently; each
request's result is bitwise equal to the same problem run alone through the
plain solver; a budget that fits one request at a time serialises them; the
footprint bounds what a live solver allocates."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _tgv(ch, n, precision="fp64"):
    prob = ch.Problem(domain=((0.0, 2 * np.pi),) * 3, bcs=((ch.PeriodicBC(), ch.PeriodicBC()),) * 3,
                      ic=lambda X, Y, Z: ch.Prim(u=(np.sin(X) * np.cos(Y), -np.cos(X) * np.sin(Y), 0 * X),
                                                 p=71.43, rho=1.0))
    return prob, ch.Numerics(n_global=(n, n, n), precision=precision)


def _shock(ch, n):
    q, L, faces = cases.shock_interface(n)
    eos = ch.IdealMixture([ch.IdealSpecies(f"s{k}", R, g) for k, (R, g) in enumerate(cases.HE_CO2)])
    bcs = ((cases.nscbc_bc(ch, faces[(0, 0)]), cases.nscbc_bc(ch, faces[(0, 1)])),
           (ch.PeriodicBC(), ch.PeriodicBC()), (ch.PeriodicBC(), ch.PeriodicBC()))
    prob = ch.Problem(domain=tuple((0.0, L[k]) for k in range(3)), bcs=bcs, ic=lambda X, Y, Z: None, eos=eos,
                      transport=ch.Transport(mu0=0.0))
    num = ch.Numerics(n_global=n, cfl=0.4)
    spec = ch.SolverSpec(prob, num)
    Q = spec.allocate_state()
    sl = tuple(slice(spec.halo[k], spec.halo[k] + n[k]) for k in range(3))
    for c, qc in enumerate(q):
        Q[sl + (c,)] = qc
    return prob, num, Q


def _alone(ctx, ch, prob, num, Q, tfinal, nmax):
    s = ch.Solver(ctx, prob, num)
    s.upload(Q)
    s.run(tfinal, nmax)
    out = s.allocate_state()
    s.download(out)
    t, st = s.t, s.step_count
    s.close()
    return out, t, st


def test_mix_equals_alone(ctx):
    import compactles_hip as ch
    reqs = []
    for n in (16, 24, 20):
        prob, num = _tgv(ch, n)
        spec = ch.SolverSpec(prob, num)
        reqs.append((prob, num, ch.initialize_(spec, spec.allocate_state(), prob.ic), 0.02, 12))
    for n in ((64, 24, 1), (48, 16, 1)):
        prob, num, Q = _shock(ch, n)
        reqs.append((prob, num, Q, 1e9, 10))
    prob, num = _tgv(ch, 16, "fp32")
    spec = ch.SolverSpec(prob, num)
    reqs.append((prob, num, ch.initialize_(spec, spec.allocate_state(), prob.ic), 0.02, 12))
    sv = ch.Server(device=0, hbm_budget=8 << 30, runners=3)
    ids = [sv.submit(p, n, tfinal=tf, nmax=nm, Q=Q) for (p, n, Q, tf, nm) in reqs]
    for rid in ids:
        assert sv.wait(rid, timeout=120.0) == ch.server.DONE
    for rid, (p, n, Q, tf, nm) in zip(ids, reqs):
        got, status, t, steps = sv.result(rid)
        ref, t_ref, st_ref = _alone(ctx, ch, p, n, Q, tf, nm)
        assert status == 0 and steps == st_ref and t == t_ref
        assert np.array_equal(got, ref), rid
    sv.close()


def test_budget_serialises_admission(ctx):
    import compactles_hip as ch
    prob, num = _tgv(ch, 24)
    fp = ch.SolverSpec(prob, num).footprint()
    sv = ch.Server(device=0, hbm_budget=int(1.5 * fp), runners=4)
    ids = [sv.submit(prob, num, tfinal=1e9, nmax=6) for _ in range(4)]
    for rid in ids:
        assert sv.wait(rid, timeout=120.0) == ch.server.DONE
    spans = sorted(sv.timing(rid)[1:3] for rid in ids)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0          # admitted only after the previous one finished
    sv.close()


def test_footprint_bounds_allocation(ctx):
    import compactles_hip as ch
    cases_ = [_tgv(ch, 32), _tgv(ch, 32, "fp32")]
    prob, num, _ = _shock(ch, (64, 24, 1))
    cases_.append((prob, num))
    for prob, num in cases_:
        s = ch.Solver(ctx, prob, num)
        used, fp = s.memory(), s.footprint()
        s.close()
        assert used <= fp <= 1.25 * used + (4 << 20), (used, fp)


def test_same_shape_requests_run_as_a_stack(ctx):
    """Same-descriptor 2-D requests queued together run as one stack (a single
    runner, so they queue); each result bitwise equal to the request alone,
    with its own end time and step cap."""
    import compactles_hip as ch
    prob, num, Q0 = _shock(ch, (64, 24, 1))
    reqs = []
    for k, (scale, tf, nm) in enumerate(((1.0, 0.02, 2 ** 62), (0.9, 0.01, 2 ** 62), (1.1, 1e9, 9), (1.0, 1e9, 4))):
        Q = Q0.copy()
        Q[..., 2:5] *= scale   # same internal energy, so every request has its own dt sequence
        Q[..., 5] += 0.5 * (scale * scale - 1) * (Q[..., 2] ** 2 + Q[..., 3] ** 2) / (scale * scale) / \
            np.maximum(Q[..., 0] + Q[..., 1], 1e-300)
        reqs.append((Q, tf, nm))
    sv = ch.Server(device=0, hbm_budget=8 << 30, runners=1)
    blocker = sv.submit(*_tgv(ch, 24), tfinal=1e9, nmax=3)   # holds the runner while the others queue
    ids = [sv.submit(prob, num, tfinal=tf, nmax=nm, Q=Q) for (Q, tf, nm) in reqs]
    for rid in [blocker] + ids:
        assert sv.wait(rid, timeout=120.0) == ch.server.DONE
    sizes = [sv.group(rid) for rid in ids]
    print(f"\nstack sizes: {sizes}")
    assert sizes == [4, 4, 4, 4]
    for rid, (Q, tf, nm) in zip(ids, reqs):
        got, status, t, steps = sv.result(rid)
        ref, t_ref, st_ref = _alone(ctx, ch, prob, num, Q, tf, nm)
        assert status == 0 and steps == st_ref and t == t_ref
        assert np.array_equal(got, ref), rid
    sv.close()
