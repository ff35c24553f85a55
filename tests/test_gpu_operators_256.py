"""The operator microbenchmark's workload as a parity test: the 256^3
periodic field of seeded Fourier modes (|k| <= 32, amplitude |k|^(-5/3),
unit RMS) through the C6 and C10 first derivatives and the C8 filter along x,
y and z, on the device, against the exact result of each periodic operator:
the inverse FFT of the field's spectrum times the operator's closed-form
modified wavenumber or transfer function (Lele 1992; Gaitonde and Visbal
2000).  Bar: FP64 1e-12 relative (infinity norm); the FP32 form on the same
input cast once is held to 2e-6 and its error printed."""
import numpy as np
import pytest

from test_oracle_operators import multiplier

pytestmark = pytest.mark.gpu
N = 256


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


@pytest.fixture(scope="module")
def field():
    from compactles_hip import workloads as W
    return W.fourier_field((N, N, N), kmax=32, seed=20261202)


def _exact(f, name, dim, h):
    ax = dim - 1
    k = np.fft.rfftfreq(N, 1.0 / N)
    m = multiplier(name, k * h)
    if name != "compact_filter":
        m = m / h
    shp = [1, 1, 1]
    shp[ax] = -1
    return np.fft.irfft(np.fft.rfft(f, axis=ax) * m.reshape(shp), n=N, axis=ax)


@pytest.mark.parametrize("name", ["lele_d1_6", "lele_d1_10", "compact_filter"])
def test_256_cube_against_closed_form(ctx, field, name):
    import compactles_hip as ch
    sch = {"lele_d1_6": ch.lele_d1_6, "lele_d1_10": ch.lele_d1_10, "compact_filter": ch.compact_filter}[name]()
    dec = ch.Decomp((N, N, N), (True, True, True), n_halo=0)
    h = 2 * np.pi / N
    f64 = ch.DeviceArray.from_numpy(ctx, field, np.float64)
    f32 = ch.DeviceArray.from_numpy(ctx, field, np.float32)
    for dim in (1, 2, 3):
        print(f"case {name} dim {dim} at {N}^3", flush=True)
        ref = _exact(field, name, dim, h)
        den = np.abs(ref).max()
        errs = {}
        for dt, f in ((np.float64, f64), (np.float32, f32)):
            out = ch.DeviceArray(ctx, (N, N, N), dt)
            ch.apply_along_(out, ch.plan_direction(ctx, dec, sch, dim, h, dtype=dt), f, dec)
            errs[np.dtype(dt).name] = np.abs(out.numpy().astype(np.float64) - ref).max() / den
            out.free()
        print(f"  {name} dim {dim}: FP64 {errs['float64']:.2e}, FP32 {errs['float32']:.2e}")
        assert errs["float64"] <= 1e-12 and errs["float32"] <= 2e-6, (name, dim, errs)
