"""Error of the FP32-mixed form against the FP64 oracle.

In the mixed form the primitives, gradients, sensors, fluxes and line solves
are FP32, while the conserved state, the LSRK register, the reduced-system
(chunk interface) correction, the conserved-state filter and every reduction
stay FP64.  The bar for kinetic energy is 1e-5 relative; the measured errors
are printed (pytest -s) so that they can be quoted."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


def _state_err(gs, os_):
    """Max over components of |GPU - oracle| / scale, where the momentum
    components share one scale (a component that is identically zero, such as
    rho w in the Taylor-Green state, is then measured against |m|)."""
    C = os_.n_cons
    Ns = C - 4
    refs = [os_.interior(os_.Q[..., c]) for c in range(C)]
    scale = [np.abs(r).max() for r in refs]
    mscale = max(scale[Ns:Ns + 3])
    for c in range(Ns, Ns + 3):
        scale[c] = mscale
    return max(np.abs(gs.field(f"Q[{c + 1}]") - refs[c]).max() / scale[c] for c in range(C))


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def test_fp32_tgv64_100_steps(ctx):
    n, L = (64, 64, 64), (TWO_PI,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    os_ = cases.oracle_solver(n, L, q, art=True)
    gs = cases.gpu_solver(ctx, n, L, q, art=True, precision="fp32")
    d1 = os_.run(1e9, 100, record=100)
    d2 = gs.run(1e9, 100, record=100)
    ke1, en1 = os_.diagnostics()
    ke2, en2 = gs.diagnostics()
    e_ke, e_en = abs(ke2 - ke1) / ke1, abs(en2 - en1) / en1
    e_dt = np.abs(d1 - d2).max() / d1.max()
    e_q = _state_err(gs, os_)
    print(f"\nFP32-mixed TGV 64^3 100 steps vs FP64 oracle: KE {e_ke:.2e}  enstrophy {e_en:.2e}  "
          f"dt {e_dt:.2e}  Q {e_q:.2e}")
    assert e_ke <= 1e-5
    assert e_en <= 1e-4
    assert e_dt <= 1e-5
    assert e_q <= 1e-4


def test_fp32_shock_interface(ctx):
    n = (96, 32, 1)
    q, L, faces = cases.shock_interface(n)
    bcs = ((1, 1), (0, 0), (0, 0))
    codes = cases.nscbc_codes(bcs, faces)
    os_ = cases.oracle_solver(n, L, q, species=cases.HE_CO2, art=True, cfl=0.4, bcs=codes, nscbc=faces)
    gs = cases.gpu_solver(ctx, n, L, q, species=cases.HE_CO2, art=True, cfl=0.4, bcs=bcs, nscbc=faces,
                          precision="fp32")
    os_.run(1e9, 40)
    gs.run(1e9, 40)
    e_q = _state_err(gs, os_)
    ke1, _ = os_.diagnostics()
    ke2, _ = gs.diagnostics()
    e_ke = abs(ke2 - ke1) / ke1
    print(f"\nFP32-mixed shock/interface 96x32 40 steps vs FP64 oracle: KE {e_ke:.2e}  Q {e_q:.2e}")
    assert e_ke <= 1e-5
    assert e_q <= 1e-4


@pytest.mark.parametrize("walls", [False, True])
def test_fp32_256_point_x_lines(ctx, walls):
    """The FP32 instances of the 256-point x-sweep kernel inside RHS, step
    and filter_state, against the FP64 oracle."""
    n, L = (256, 16, 12), (TWO_PI, 0.5, 0.4)
    per = (not walls, True, True)
    X, Y, Z = cases.grid(n, L, periodic=per)
    ky, kz = 2 * np.pi / L[1], 2 * np.pi / L[2]
    prim = (0.2 * np.sin(X) * np.cos(ky * Y) + (0.1 * np.sin(0.5 * X) ** 2 if walls else 0),
            0.05 * np.cos(2 * X + ky * Y), 0.04 * np.sin(X + kz * Z),
            1.0 + 0.1 * np.cos(X) * np.cos(kz * Z), 1.0 + 0.05 * np.sin(ky * Y + X))
    q = cases.conserved(prim)
    bcs = ((1, 1), (0, 0), (0, 0)) if walls else None
    os_ = cases.oracle_solver(n, L, q, mu0=1e-3, art=True, bcs=bcs)
    gs = cases.gpu_solver(ctx, n, L, q, mu0=1e-3, art=True, bcs=bcs, precision="fp32", chunks=(16, 32, 32))
    os_.run(1e9, 5)
    gs.run(1e9, 5)
    e_q = _state_err(gs, os_)
    print(f"\nFP32-mixed 256x16x12 {'walls' if walls else 'periodic'}, 5 steps vs FP64 oracle: Q {e_q:.2e}")
    assert e_q <= 1e-5
