"""GPU parity of the operator plugin (cles_apply_along through the C ABI)
against the oracle's apply_along! on the same seeded inputs.  Bar: FP64
normwise relative error <= 1e-12 per operator; FP32 <= 2e-5 against the FP64
oracle (inputs cast once)."""
import numpy as np
import pytest

import helpers as Hh
import oracle as O

pytestmark = pytest.mark.gpu

SCHEMES = ["lele_d1_6", "pade_d1_4", "lele_d1_10", "compact_filter"]
GRIDS = [(24, 20, 16), (64, 40, 36), (100, 33, 70), (256, 12, 10)]


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _py_scheme(name):
    import compactles_hip as ch
    return {"lele_d1_6": ch.lele_d1_6, "pade_d1_4": ch.pade_d1_4, "lele_d1_10": ch.lele_d1_10,
            "compact_filter": ch.compact_filter}[name]()


def _oracle_apply(name, dim, n, periodic, h, f_int):
    plan = O.Plan(O.scheme(name), dim, n, periodic, h)
    F = Hh.pad(f_int, n)
    O.wrap_halos(F, n, Hh.halo_for(n), (periodic,) * 3)
    return Hh.interior(plan.apply(F), n)


def _gpu_apply(ctx, name, dim, n, periodic, h, f_int, dtype=np.float64, padded=True):
    import compactles_hip as ch
    dec = ch.Decomp(tuple(n), (periodic,) * 3, n_halo=4 if padded else 0)
    plan = ch.plan_direction(ctx, dec, _py_scheme(name), dim, h, dtype=dtype)
    if padded:
        fh = Hh.pad(f_int, n)
    else:
        fh = np.asfortranarray(f_int)
    f = ch.DeviceArray.from_numpy(ctx, fh, dtype)
    out = ch.field(ctx, dec, dtype)
    ch.apply_along_(out, plan, f, dec)
    res = out.numpy().astype(np.float64)
    return Hh.interior(res, n) if padded else res


@pytest.mark.parametrize("name", SCHEMES)
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("grid", GRIDS)
def test_fp64_matches_oracle(ctx, name, periodic, grid):
    n = grid
    f = Hh.random_field(n, seed=11)
    for dim in (1, 2, 3):
        h = 0.37 / n[dim - 1]
        ref = _oracle_apply(name, dim, n, periodic, h, f)
        got = _gpu_apply(ctx, name, dim, n, periodic, h, f)
        err = Hh.relerr(got, ref)
        assert err <= 1e-12, (name, dim, periodic, n, err)


@pytest.mark.parametrize("name", SCHEMES)
def test_unpadded_layout_equals_padded(ctx, name):
    n = (48, 40, 36)
    f = Hh.random_field(n, seed=3)
    for dim in (1, 2, 3):
        a = _gpu_apply(ctx, name, dim, n, True, 0.1, f, padded=True)
        b = _gpu_apply(ctx, name, dim, n, True, 0.1, f, padded=False)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", SCHEMES)
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("n", [(64, 48, 40), (256, 12, 10), (24, 20, 16)])
def test_fp32_against_fp64_oracle(ctx, name, periodic, n):
    f = Hh.random_field(n, seed=5)
    f32 = f.astype(np.float32).astype(np.float64)
    for dim in (1, 2, 3):
        h = 2 * np.pi / n[dim - 1]
        ref = _oracle_apply(name, dim, n, periodic, h, f32)
        got = _gpu_apply(ctx, name, dim, n, periodic, h, f, dtype=np.float32)
        assert Hh.relerr(got, ref) <= 2e-5, (name, dim, periodic)


def test_in_place_filter(ctx):
    import compactles_hip as ch
    n = (64, 64, 64)
    f = Hh.random_field(n, seed=9)
    dec = ch.Decomp(n, (True, True, True))
    plans = [ch.plan_direction(ctx, dec, ch.compact_filter(), d, 1.0) for d in (1, 2, 3)]
    a = ch.DeviceArray.from_numpy(ctx, Hh.pad(f, n))
    ch.filter_field_(a, plans, dec)
    got = Hh.interior(a.numpy(), n)
    ref = f
    for d in (1, 2, 3):
        ref = _oracle_apply("compact_filter", d, n, True, 1.0, ref)
    assert Hh.relerr(got, ref) <= 1e-12


def test_too_small_extent_is_an_error(ctx):
    import compactles_hip as ch
    dec = ch.Decomp((8, 16, 16), (False, True, True))
    with pytest.raises(ch.ClesError) as ei:
        ch.plan_direction(ctx, dec, ch.compact_filter(), 1, 0.1)
    assert ei.value.code == 2 and "too small" in str(ei.value)


# ---- parity-folded plans: plan_direction(...; lo_fold, hi_fold) -------------
def _oracle_fold_apply(name, dim, n, h, f_int, lo, hi):
    """The reference's route: mirror-fill the halos with the input's sign
    (fold_fill!, sigma_f = sigma_g for a filter, -sigma_g for a derivative;
    folds.jl:271-281), then apply the folded plan."""
    plan = O.Plan(O.scheme(name), dim, n, False, h, lo_fold=lo, hi_fold=hi)
    sg = lo or hi
    sf = sg if name == "compact_filter" else -sg
    F = Hh.pad(f_int, n)
    O.fold_fill(F, n, Hh.halo_for(n), dim, bool(lo), bool(hi), sf)
    return Hh.interior(plan.apply(F), n)


def _gpu_fold_apply(ctx, name, dim, n, h, f_int, lo, hi, dtype=np.float64):
    """The device plan reads no halo: they hold NaN here."""
    import compactles_hip as ch
    dec = ch.Decomp(tuple(n), (False,) * 3)
    plan = ch.plan_direction(ctx, dec, _py_scheme(name), dim, h, dtype=dtype, lo_fold=lo or None,
                             hi_fold=hi or None)
    fh = np.full(Hh.padded_shape(n), np.nan, order="F")
    hl = Hh.halo_for(n)
    fh[hl[0]:hl[0] + n[0], hl[1]:hl[1] + n[1], hl[2]:hl[2] + n[2]] = f_int
    f = ch.DeviceArray.from_numpy(ctx, fh, dtype)
    out = ch.field(ctx, dec, dtype)
    ch.apply_along_(out, plan, f, dec)
    return Hh.interior(out.numpy().astype(np.float64), n)


@pytest.mark.parametrize("name", SCHEMES)
@pytest.mark.parametrize("ends", [(1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("sigma_g", [1, -1])
@pytest.mark.parametrize("grid", [(24, 20, 16), (100, 33, 70), (256, 12, 40)])
def test_folded_plans_match_oracle(ctx, name, ends, sigma_g, grid):
    n = grid
    f = Hh.random_field(n, seed=13)
    lo, hi = sigma_g * ends[0], sigma_g * ends[1]
    for dim in (1, 2, 3):
        h = 0.29 / n[dim - 1]
        ref = _oracle_fold_apply(name, dim, n, h, f, lo, hi)
        got = _gpu_fold_apply(ctx, name, dim, n, h, f, lo, hi)
        err = Hh.relerr(got, ref)
        assert np.all(np.isfinite(got)) and err <= 1e-12, (name, ends, sigma_g, dim, n, err)


@pytest.mark.parametrize("name", SCHEMES)
def test_folded_plans_fp32(ctx, name):
    n = (64, 48, 40)
    f = Hh.random_field(n, seed=17)
    f32 = f.astype(np.float32).astype(np.float64)
    for dim in (1, 2, 3):
        h = 0.29 / n[dim - 1]
        for lo, hi in ((1, 0), (-1, -1)):
            ref = _oracle_fold_apply(name, dim, n, h, f32, lo, hi)
            got = _gpu_fold_apply(ctx, name, dim, n, h, f, lo, hi, dtype=np.float32)
            assert Hh.relerr(got, ref) <= 2e-5, (name, dim, lo, hi)


def test_fold_keywords_are_checked(ctx):
    import compactles_hip as ch
    dec = ch.Decomp((24, 20, 16), (False, True, True))
    with pytest.raises(ValueError):
        ch.plan_direction(ctx, dec, ch.lele_d1_6(), 1, 0.1, lo_fold=2)
    # periodic dimension: no physical edge, the keywords change nothing
    a = ch.plan_direction(ctx, dec, ch.lele_d1_6(), 2, 0.1, lo_fold=1, hi_fold=1)
    f = Hh.random_field((24, 20, 16), seed=1)
    got = _gpu_apply(ctx, "lele_d1_6", 2, (24, 20, 16), True, 0.1, f)
    fd = ch.DeviceArray.from_numpy(ctx, Hh.pad(f, (24, 20, 16)))
    dec2 = ch.Decomp((24, 20, 16), (True, True, True))
    out = ch.field(ctx, dec2)
    ch.apply_along_(out, a, fd, dec2)
    assert np.array_equal(Hh.interior(out.numpy(), (24, 20, 16)), got)
