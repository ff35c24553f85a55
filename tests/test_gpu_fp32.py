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
