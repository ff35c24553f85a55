"""Pins the oracle's compact operators (CPU, no GPU needed).

Independent checks, each able to catch a dropped term, a wrong sign or index:
  * periodic operators against the closed-form modified wavenumber / transfer
    function written from the published scheme parameters (Lele 1992: a, b, c,
    alpha, beta; Gaitonde-Visbal filter table), applied by FFT;
  * closed-edge operators against a dense solve of the matrix assembled from
    the scheme definition (numpy.linalg.solve);
  * the reference's own tolerance tests (test/runtests.jl line ranges cited).
"""
import numpy as np
import pytest

import helpers as Hh
import oracle as O

TWO_PI = 2 * np.pi


def apply(name, dim, n, periodic, h, f_int, alphaf=0.45):
    plan = O.Plan(O.scheme(name, alphaf), dim, n, periodic, h)
    F = Hh.pad(f_int, n)
    O.wrap_halos(F, n, Hh.halo_for(n), (periodic,) * 3)
    return Hh.interior(plan.apply(F), n)


# ---- closed forms (Lele 1992 eqs. 2.1-2.2; Gaitonde & Visbal 2000) ---------
def multiplier(name, w, alphaf=0.45):
    if name == "lele_d1_6":          # alpha = 1/3, a = 14/9, b = 1/9
        a, b, al = 14 / 9, 1 / 9, 1 / 3
        return 1j * (a * np.sin(w) + b / 2 * np.sin(2 * w)) / (1 + 2 * al * np.cos(w))
    if name == "pade_d1_4":          # alpha = 1/4, a = 3/2
        return 1j * (1.5 * np.sin(w)) / (1 + 0.5 * np.cos(w))
    if name == "lele_d1_10":         # alpha = 1/2, beta = 1/20, a = 17/12, b = 101/150, c = 1/100
        a, b, c, al, be = 17 / 12, 101 / 150, 1 / 100, 0.5, 0.05
        num = a * np.sin(w) + b / 2 * np.sin(2 * w) + c / 3 * np.sin(3 * w)
        return 1j * num / (1 + 2 * al * np.cos(w) + 2 * be * np.cos(2 * w))
    af = alphaf                      # eighth-order filter: sum_n a_n cos(n w) / (1 + 2 af cos w)
    an = [(93 + 70 * af) / 128, (7 + 18 * af) / 16, (-7 + 14 * af) / 32, (1 - 2 * af) / 16, (-1 + 2 * af) / 128]
    return (sum(an[m] * np.cos(m * w) for m in range(5))) / (1 + 2 * af * np.cos(w))


@pytest.mark.parametrize("name", ["lele_d1_6", "pade_d1_4", "lele_d1_10", "compact_filter"])
def test_periodic_closed_form(name):
    n = (32, 24, 20)
    f = Hh.fourier_field(n, kmax=7, seed=20261202)
    for dim in (1, 2, 3):
        h = TWO_PI / n[dim - 1]
        got = apply(name, dim, n, True, h, f)
        k = np.fft.fftfreq(n[dim - 1], 1.0 / n[dim - 1])
        mult = multiplier(name, k * h)
        if name != "compact_filter":
            mult = mult / h
        shp = [1, 1, 1]
        shp[dim - 1] = -1
        ref = np.real(np.fft.ifft(np.fft.fft(f, axis=dim - 1) * mult.reshape(shp), axis=dim - 1))
        assert Hh.relerr(got, ref) <= 1e-13, (name, dim)


def test_filter_closed_form_other_strength():
    n = (40, 8, 8)
    f = Hh.random_field(n, seed=2)
    got = apply("compact_filter", 1, n, True, 1.0, f, alphaf=0.3)
    w = np.fft.fftfreq(40, 1 / 40) * TWO_PI / 40
    ref = np.real(np.fft.ifft(np.fft.fft(f, axis=0) * multiplier("compact_filter", w, 0.3)[:, None, None], axis=0))
    assert Hh.relerr(got, ref) <= 1e-13


# ---- closed edges: dense matrix from the scheme definition -----------------
SCHEMES = {
    # name: (q, lhs, a0, coeffs, symmetric, closures[(lhs_centred, rhs_from_edge)])
    "lele_d1_6": (1, [1 / 3], 0.0, [7 / 9, 1 / 36], False,
                  [([0, 1, 2], [-5 / 2, 2, 1 / 2]), ([1 / 4, 1, 1 / 4], [-3 / 4, 0, 3 / 4])]),
    "lele_d1_10": (2, [1 / 2, 1 / 20], 0.0, [17 / 24, 101 / 600, 1 / 600], False,
                   [([0, 0, 1, 2, 0], [-5 / 2, 2, 1 / 2]), ([0, 1 / 4, 1, 1 / 4, 0], [-3 / 4, 0, 3 / 4]),
                    ([0, 1 / 3, 1, 1 / 3, 0], [-1 / 36, -7 / 9, 0, 7 / 9, 1 / 36])]),
}


def dense_operator(q, lhs, a0, coeffs, sym, closures, n, h):
    """A g = B f for one closed line: interior rows, low closure rows as given,
    high rows mirrored (LHS reversed, RHS reversed and negated for derivatives),
    derivative RHS scaled by 1/h (src/kernels.jl:1-16 semantics)."""
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    scale = 1.0 if sym else 1.0 / h
    for i in range(n):
        A[i, i] = 1.0
        for s, v in enumerate(lhs, start=1):
            if i - s >= 0:
                A[i, i - s] = v
            if i + s < n:
                A[i, i + s] = v
        B[i, i] = a0 * scale if sym else 0.0
        for m, c in enumerate(coeffs, start=1):
            if 0 <= i + m < n:
                B[i, i + m] += c * scale
            if 0 <= i - m < n:
                B[i, i - m] += (c if sym else -c) * scale
    for j, (cl, rhs) in enumerate(closures):
        for r in (j, n - 1 - j):
            A[r, :] = 0.0
            B[r, :] = 0.0
        for s in range(-q, q + 1):
            if 0 <= j + s < n:
                A[j, j + s] = cl[q + s]
            if 0 <= n - 1 - j - s < n:
                A[n - 1 - j, n - 1 - j - s] = cl[q + s]
        for kk, v in enumerate(rhs):
            B[j, kk] = v * scale
            B[n - 1 - j, n - 1 - kk] = v * scale * (1 if sym else -1)
    return A, B


@pytest.mark.parametrize("name", ["lele_d1_6", "lele_d1_10"])
def test_closed_edges_dense_solve(name):
    q, lhs, a0, coeffs, sym, closures = SCHEMES[name]
    n = (37, 9, 8)
    f = Hh.random_field(n, seed=7)
    for dim in (1, 2, 3):
        nd = n[dim - 1]
        if nd < 2 * len(closures) + 1:
            continue
        h = 0.1
        A, B = dense_operator(q, lhs, a0, coeffs, sym, closures, nd, h)
        M = np.linalg.solve(A, B)
        ref = np.moveaxis(np.tensordot(M, np.moveaxis(f, dim - 1, 0), axes=1), 0, dim - 1)
        got = apply(name, dim, n, False, h, f)
        assert Hh.relerr(got, ref) <= 1e-12, (name, dim)


def test_filter_closed_dense_solve():
    af = 0.45
    b2 = ((1 + 2 * af) / 2, (1 + 2 * af) / 2)
    b4 = ((5 + 6 * af) / 8, (1 + 2 * af) / 2, (-1 + 2 * af) / 8)
    b6 = ((11 + 10 * af) / 16, (15 + 34 * af) / 32, (-3 + 6 * af) / 16, (1 - 2 * af) / 32)

    def centred(b):
        return [b[m] / 2 for m in range(len(b) - 1, 0, -1)] + [b[0]] + [b[m] / 2 for m in range(1, len(b))]

    an = [(93 + 70 * af) / 128, (7 + 18 * af) / 16, (-7 + 14 * af) / 32, (1 - 2 * af) / 16, (-1 + 2 * af) / 128]
    closures = [([0, 1, 0], [1.0]), ([af, 1, af], centred(b2)), ([af, 1, af], centred(b4)),
                ([af, 1, af], centred(b6))]
    nd = 29
    A, B = dense_operator(1, [af], an[0], [a / 2 for a in an[1:]], True, closures, nd, 1.0)
    n = (nd, 10, 9)
    f = Hh.random_field(n, seed=8)
    ref = np.tensordot(np.linalg.solve(A, B), f, axes=1)
    got = apply("compact_filter", 1, n, False, 1.0, f)
    assert Hh.relerr(got, ref) <= 1e-12
    # consistency of every closure row: RHS weights sum to 1 + 2 af (kernels.jl:98)
    for _, rhs in closures[1:]:
        assert abs(sum(rhs) - (1 + 2 * af)) < 1e-14


# ---- the reference's own operator tests -----------------------------------
def Hh_grid(n, L, periodic):
    xs = []
    for d in range(3):
        h = L[d] / (n[d] if periodic[d] else n[d] - 1)
        xs.append(np.arange(n[d]) * h)
    return np.meshgrid(*xs, indexing="ij"), [L[d] / (n[d] if periodic[d] else n[d] - 1) for d in range(3)]


def test_c6_spectral_accuracy():
    """test/runtests.jl:86-97."""
    n = (32, 32, 32)
    (X, Y, Z), h = Hh_grid(n, (TWO_PI,) * 3, (True,) * 3)
    f = np.sin(3 * X) * np.cos(2 * Y)
    assert np.abs(apply("lele_d1_6", 1, n, True, h[0], f) - 3 * np.cos(3 * X) * np.cos(2 * Y)).max() < 1e-4
    assert np.abs(apply("lele_d1_6", 2, n, True, h[1], f) + 2 * np.sin(3 * X) * np.sin(2 * Y)).max() < 2e-5
    assert np.abs(apply("lele_d1_6", 3, n, True, h[2], f)).max() < 1e-10


def test_c10_accuracy_and_closures():
    """test/runtests.jl:99-132."""
    n = (32, 32, 32)
    (X, Y, Z), h = Hh_grid(n, (TWO_PI,) * 3, (True,) * 3)
    f = np.sin(3 * X) * np.cos(2 * Y)
    e10 = np.abs(apply("lele_d1_10", 1, n, True, h[0], f) - 3 * np.cos(3 * X) * np.cos(2 * Y)).max()
    assert e10 < 1e-7
    assert np.abs(apply("lele_d1_10", 2, n, True, h[1], f) + 2 * np.sin(3 * X) * np.sin(2 * Y)).max() < 1e-8
    assert np.abs(apply("lele_d1_10", 3, n, True, h[2], f)).max() < 1e-12
    e6 = np.abs(apply("lele_d1_6", 1, n, True, h[0], f) - 3 * np.cos(3 * X) * np.cos(2 * Y)).max()
    assert e10 < e6 / 100
    nc = (32, 12, 12)
    (X, Y, Z), h = Hh_grid(nc, (1.0, 1.0, 1.0), (False, True, True))
    g = 1 + 2 * X + 3 * X ** 2 - X ** 3
    assert np.abs(apply("lele_d1_10", 1, nc, False, h[0], g) - (2 + 6 * X - 3 * X ** 2)).max() < 1e-10


def test_pade_fourth_order():
    """test/runtests.jl:134-149."""
    errs = []
    for N in (16, 32, 64):
        n = (N, 12, 12)
        (X, Y, Z), h = Hh_grid(n, (TWO_PI,) * 3, (True,) * 3)
        errs.append(np.abs(apply("pade_d1_4", 1, n, True, h[0], np.sin(X)) - np.cos(X)).max())
    p = np.log(errs[0] / errs[2]) / np.log(4)
    assert 3.5 < p < 4.5 and errs[2] < errs[1] < errs[0]


@pytest.mark.parametrize("name,expect", [("lele_d1_6", (5.8, 6.3)), ("lele_d1_10", (9.5, 10.6))])
def test_interior_order(name, expect):
    """test/convergence.jl:20-25 (measured orders C6 6.01, C10 10.04)."""
    errs = []
    for N in (16, 32):
        n = (N, 8, 8)
        (X, Y, Z), h = Hh_grid(n, (TWO_PI,) * 3, (True,) * 3)
        errs.append(np.abs(apply(name, 1, n, True, h[0], np.sin(X)) - np.cos(X)).max())
    p = np.log(errs[0] / errs[1]) / np.log(2)
    assert expect[0] < p < expect[1]


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_closures_polynomial_exactness(dim):
    """test/runtests.jl:151-169 and :205-214: deg-3 polynomials differentiate
    exactly through the C6 closure rows on every sweep direction."""
    n = tuple(32 if d == dim - 1 else 12 for d in range(3))
    per = tuple(d != dim - 1 for d in range(3))
    (X, Y, Z), h = Hh_grid(n, (1.0, 1.0, 1.0), per)
    t = (X, Y, Z)[dim - 1]
    got = apply("lele_d1_6", dim, n, False, h[dim - 1], 1 + 2 * t + 3 * t ** 2 - t ** 3)
    assert np.abs(got - (2 + 6 * t - 3 * t ** 2)).max() < 1e-10


def test_transposed_path_equals_x_path():
    """test/runtests.jl:190-203."""
    n = (24, 24, 24)
    (X, Y, Z), h = Hh_grid(n, (TWO_PI,) * 3, (True,) * 3)

    def fn(x, y):
        return np.sin(2 * x + 0.3) * np.cos(y) + 0.1

    d1 = apply("lele_d1_6", 1, n, True, h[0], fn(X, Y))
    d2 = apply("lele_d1_6", 2, n, True, h[1], fn(Y, X))
    assert np.abs(d1 - np.transpose(d2, (1, 0, 2))).max() < 1e-11


def test_filter_constants_exact_nyquist_damped():
    """test/runtests.jl:216-229 (periodic 32 x 12 x 12)."""
    n = (32, 12, 12)
    got = np.ones(n)
    for d in (1, 2, 3):
        got = apply("compact_filter", d, n, True, 1.0, got)
    assert np.abs(got - 1).max() < 1e-12
    got = 1.0 + 0.5 * (-1.0) ** np.arange(1, 33)[:, None, None] * np.ones(n)
    for d in (1, 2, 3):
        got = apply("compact_filter", d, n, True, 1.0, got)
    assert np.abs(got[:, 0, 0] - 1.0).max() < 0.35


def test_filter_constants_exact_through_closures():
    n = (32, 12, 12)
    got = apply("compact_filter", 1, n, False, 1.0, 3.5 * np.ones(n))
    assert np.abs(got - 3.5).max() < 1e-12


def test_extent_validation_messages():
    """operators.jl:50-53: extent too small for the scheme."""
    with pytest.raises(ValueError, match="too small"):
        O.Plan(O.scheme("compact_filter"), 1, (8, 12, 12), False, 1.0)


# ---- parity-folded plans (plan_direction's lo_fold / hi_fold) ---------------
# A fold at a half-offset edge is the same operator as the unfolded one on the
# line doubled by the mirror: the doubled data has parity sigma_f, so by
# uniqueness its solution has parity sigma_g (= sigma_f for a filter, -sigma_f
# for a derivative, folds.jl:271-274), and the rows at the mirror reduce to the
# folded rows of operators.jl:59-70 / operators_banded.jl:55-73.  One folded
# end doubles onto a line closed at both ends; two folded ends of the same sign
# double onto a periodic line of 2n points.
def fold_sigma_f(name, sigma_g):
    return sigma_g if name == "compact_filter" else -sigma_g


def folded_apply(name, dim, n, h, f_int, lo, hi):
    """The oracle's folded plan on mirror-filled halos, as fold_apply! runs it."""
    plan = O.Plan(O.scheme(name), dim, n, False, h, lo_fold=lo, hi_fold=hi)
    sf = fold_sigma_f(name, lo or hi)
    F = Hh.pad(f_int, n)
    O.fold_fill(F, n, Hh.halo_for(n), dim, bool(lo), bool(hi), sf)
    return Hh.interior(plan.apply(F), n)


def doubled_apply(name, dim, n, h, f_int, lo, hi):
    ax = dim - 1
    sf = fold_sigma_f(name, lo or hi)
    mirror = sf * np.flip(f_int, axis=ax)
    if lo and hi:
        ext, periodic, keep = np.concatenate([f_int, mirror], axis=ax), True, slice(0, n[ax])
    elif lo:
        ext, periodic, keep = np.concatenate([mirror, f_int], axis=ax), False, slice(n[ax], 2 * n[ax])
    else:
        ext, periodic, keep = np.concatenate([f_int, mirror], axis=ax), False, slice(0, n[ax])
    n2 = list(n)
    n2[ax] *= 2
    got = apply(name, dim, tuple(n2), periodic, h, np.asfortranarray(ext))
    idx = [slice(None)] * 3
    idx[ax] = keep
    return got[tuple(idx)]


@pytest.mark.parametrize("name", ["lele_d1_6", "lele_d1_10", "pade_d1_4", "compact_filter"])
@pytest.mark.parametrize("ends", [(1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("sigma_g", [1, -1])
def test_folded_plan_equals_doubled_line(name, ends, sigma_g):
    n = (18, 13, 11)
    f = Hh.random_field(n, seed=7)
    for dim in (1, 2, 3):
        h = 0.37 / n[dim - 1]
        lo, hi = sigma_g * ends[0], sigma_g * ends[1]
        got = folded_apply(name, dim, n, h, f, lo, hi)
        ref = doubled_apply(name, dim, n, h, f, lo, hi)
        assert Hh.relerr(got, ref) <= 1e-13, (name, ends, sigma_g, dim, Hh.relerr(got, ref))


def test_folded_plan_wrong_parity_is_caught():
    """The doubled-line check separates the two parities: the plan of the other
    sign differs at O(1) near the fold."""
    n = (18, 13, 11)
    f = Hh.random_field(n, seed=7)
    h = 0.37 / 18
    right = doubled_apply("lele_d1_6", 1, n, h, f, 1, 0)
    plan = O.Plan(O.scheme("lele_d1_6"), 1, n, False, h, lo_fold=-1)
    F = Hh.pad(f, n)
    O.fold_fill(F, n, Hh.halo_for(n), 1, True, False, -1)
    wrong = Hh.interior(plan.apply(F), n)
    assert Hh.relerr(wrong, right) > 1e-2


def test_fold_on_a_periodic_line_is_ignored_by_the_oracle():
    """One rank of a periodic dimension has no physical edge (at_lo_edge is
    false), so the keywords do nothing there (operators.jl:57-58)."""
    n = (16, 9, 9)
    f = Hh.fourier_field(n, kmax=3, seed=3)
    h = TWO_PI / 16
    a = O.Plan(O.scheme("lele_d1_6"), 1, n, True, h, lo_fold=1, hi_fold=1)
    b = O.Plan(O.scheme("lele_d1_6"), 1, n, True, h)
    F = Hh.pad(f, n)
    O.wrap_halos(F, n, Hh.halo_for(n), (True,) * 3)
    assert np.array_equal(a.apply(F), b.apply(F))
