"""The C ABI boundary, without a GPU: the library loads, exports every
function include/compactles_hip.h declares, the ctypes records match the C
layout, and the two independent restatements of the scheme presets (the
product's Python mirror and the oracle) agree field by field."""
import ctypes as C
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "compactles_hip.h")


def header_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void \*)\s*(cles_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    import compactles_hip as ch
    lib = ch.load()
    declared = header_functions()
    assert len(declared) >= 30
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    from compactles_hip import _lib
    assert sorted(_lib.SIGNATURES) == declared


def test_abi_version():
    import compactles_hip as ch
    assert ch.load().cles_abi_version() == 5


def test_struct_layout_matches_header():
    """sizeof/offsetof from the C compiler against the ctypes records."""
    from compactles_hip import _lib as L
    from compactles_hip.callbacks import RunHooks
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "compactles_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(cles_scheme_desc), sizeof(cles_nscbc_desc), sizeof(cles_solver_desc),
         offsetof(cles_solver_desc, precision));
  printf("%zu %zu %zu %zu\n", offsetof(cles_scheme_desc, closure_rhs), offsetof(cles_solver_desc, filt),
         offsetof(cles_solver_desc, bc_state), offsetof(cles_solver_desc, nscbc));
  printf("%zu %zu %zu %zu\n", offsetof(cles_solver_desc, retries), offsetof(cles_solver_desc, cfl_backoff),
         offsetof(cles_solver_desc, landing_steps), sizeof(cles_run_hooks));
  printf("%zu %zu\n", offsetof(cles_run_hooks, after_step), offsetof(cles_run_hooks, rewind));
  printf("%zu %zu %zu %zu\n", offsetof(cles_solver_desc, eos_kind), offsetof(cles_solver_desc, T_guess),
         offsetof(cles_solver_desc, nasa9), sizeof(cles_nasa9_species));
  printf("%zu %zu %zu\n", offsetof(cles_solver_desc, metric), offsetof(cles_solver_desc, Twall),
         offsetof(cles_solver_desc, bc_after));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        exe = os.path.join(td, "p")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [C.sizeof(L.SchemeDesc), C.sizeof(L.NscbcDesc), C.sizeof(L.SolverDesc), L.SolverDesc.precision.offset,
            L.SchemeDesc.closure_rhs.offset, L.SolverDesc.filt.offset, L.SolverDesc.bc_state.offset,
            L.SolverDesc.nscbc.offset, L.SolverDesc.retries.offset, L.SolverDesc.cfl_backoff.offset,
            L.SolverDesc.landing_steps.offset, C.sizeof(RunHooks), RunHooks.after_step.offset, RunHooks.rewind.offset,
            L.SolverDesc.eos_kind.offset, L.SolverDesc.T_guess.offset, L.SolverDesc.nasa9.offset,
            C.sizeof(L.Nasa9SpeciesDesc), L.SolverDesc.metric.offset, L.SolverDesc.Twall.offset,
            L.SolverDesc.bc_after.offset]
    assert got == want


@pytest.mark.parametrize("name", ["lele_d1_6", "pade_d1_4", "lele_d1_10", "compact_filter"])
def test_scheme_presets_agree_with_oracle(name):
    import compactles_hip as ch
    import oracle as O
    py = getattr(ch, name)().to_desc()
    orc = O.scheme(name)
    assert py.q == orc.q and py.M == orc.M and py.symmetric == orc.symmetric and py.nclosure == orc.nclosure
    assert py.a0 == orc.a0
    assert list(py.lhs) == list(orc.lhs)
    assert list(py.coeffs) == list(orc.coeffs)
    for j in range(py.nclosure):
        w = 2 * py.q + 1
        assert list(py.closure_lhs[j])[:w] == list(orc.cl_lhs[j])[:w]
        assert py.closure_rhs_len[j] == orc.cl_rhs_len[j]
        assert list(py.closure_rhs[j]) == list(orc.cl_rhs[j])


def test_no_device_is_a_loud_error():
    """Without a gfx950 device the context cannot be created (no host fallback)."""
    import compactles_hip as ch
    code = "import sys; sys.path.insert(0, %r)\nimport compactles_hip as ch\n" \
           "try:\n    ch.Context(0)\nexcept ch.ClesError as e:\n    print('ERR', e.code)\nelse:\n    print('OK')\n" % \
           os.path.join(ROOT, "compactles.jl_amd")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    has_gpu = os.path.exists("/dev/kfd")
    if not has_gpu:
        assert "ERR" in r.stdout, (r.stdout, r.stderr)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_python_frontend_validation():
    """problem.jl:49-57 Prim validation and conserved_from_prim."""
    import compactles_hip as ch
    with pytest.raises(ValueError):
        ch.Prim(p=1.0, rho=1.0, T_ion=1.0)
    with pytest.raises(ValueError):
        ch.Prim(p=1.0, rho=1.0, Y=(0.5, 0.4))
    eos = ch.single_species(gamma=1.4)
    q = ch.conserved_from_prim(eos, ch.Prim(u=(1.0, 2.0, 3.0), p=0.4, rho=2.0))
    # rho E = p/(gamma-1) + rho |u|^2 / 2
    assert np.allclose(q, [2.0, 2.0, 4.0, 6.0, 0.4 / 0.4 + 0.5 * 2.0 * 14.0], rtol=1e-15)


def test_dirichlet_face_state_must_be_constant():
    """The device holds one Dirichlet state per face: a state that varies over
    the face or in time is refused at setup (problem.jl:192-207 evaluates
    fun(x, t) at every stage time)."""
    import compactles_hip as ch
    from compactles_hip.problem import SolverSpec

    def prob(fun):
        bcs = ((ch.DirichletBC(fun), ch.DirichletBC(fun)), (ch.PeriodicBC(), ch.PeriodicBC()),
               (ch.PeriodicBC(), ch.PeriodicBC()))
        return ch.Problem(domain=((0.0, 1.0),) * 3, bcs=bcs,
                          ic=lambda x, y, z: ch.Prim(p=1.0 + 0 * x, rho=1.0 + 0 * x))

    num = ch.Numerics(n_global=(16, 8, 1))
    ok = SolverSpec(prob(lambda x, y, z, t: ch.Prim(u=(0.5, 0.0, 0.0), p=1.0, rho=1.0)), num)
    assert ok.desc.bc_state[0][0][1] == 0.5
    with pytest.raises(ValueError, match="uniform over the face"):
        SolverSpec(prob(lambda x, y, z, t: ch.Prim(p=1.0 + y, rho=1.0 + 0 * y)), num)
    with pytest.raises(ValueError, match="constant in time"):
        SolverSpec(prob(lambda x, y, z, t: ch.Prim(p=1.0 + t, rho=1.0)), num)


def test_bench_traffic_lookup_skips_earlier_build_entries():
    # counter lines kept from an earlier build sit under another key and must
    # not be read as the current kernel's traffic
    sys.path.insert(0, ROOT)
    import bench
    assert bench.read_traffic(256, "fp64", "cles::k_fluxes<double, true, true>") is None
    assert bench.read_traffic(256, "fp64", "cles::k_strain<double>") > 0
    assert bench.read_traffic(256, "fp32", bench.DOMINANT_KERNEL["fp32"]) > 0
