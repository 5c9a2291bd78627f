"""GPU: the reduced-precision Poisson solves against the FP64 solve (FP64 CG
with the FP64 multigrid cycle) on the 60x120x36-cell FinFET mesh with a real
right-hand side:

- the mixed solve (FP32 bulk with FP64 iterative refinement,
  emc_poisson_solve_mixed);
- the FP64 CG preconditioned by the FP32 cycle (the library's default).

Each meets ||r|| <= 1e-10 ||b||, the residual checked on the operator as
assembled by the oracle (init_matrix with the contact rows), and the potentials
agree to the accuracy that tolerance allows (stated below); step loops with
each solve run and keep the FP64 tolerance."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from helpers import cfg1_deck, emcgpu

pytestmark = pytest.mark.gpu


def test_mixed_solve_meets_the_fp64_tolerance():
    eng = emcgpu.Engine(cfg1_deck(refine=3, electrons=60000, holes=6000), device=0, poisson_fp32=0)
    eng.particles_init()
    node_q, _, _, _ = eng.deposit()
    rhs = eng.rhs_from_charge(node_q)
    print("[gpu] FP64 and mixed solves, 60x120x36", flush=True)
    x64, it64, rel64 = eng.poisson_solve(rhs, rtol=1e-10)
    xm, outer, inner, relm = eng.poisson_solve_mixed(rhs, rtol=1e-10, inner_rtol=1e-5)
    assert rel64 <= 1e-10 and relm <= 1e-10
    assert 1 <= outer <= 4 and inner > 0
    lib = oracle_lib.load()
    rhs_ref = np.zeros_like(rhs)
    lib.orc_rhs(C.byref(eng.mesh), oracle_lib.i64ptr(node_q), None, oracle_lib.dptr(rhs_ref), None)
    ax = np.zeros_like(rhs)
    lib.orc_apply_full(C.byref(eng.mesh), oracle_lib.dptr(xm), oracle_lib.dptr(ax))
    assert np.linalg.norm(ax - rhs_ref) <= 1e-9 * np.linalg.norm(rhs_ref)
    # two solutions of the same system to 1e-10 in the residual: their difference
    # is the error of each, bounded by the tolerance times the condition number
    # (about 1e3 for this operator); in volts, far below a microvolt
    pot0 = eng.mesh.pot0
    err = np.max(np.abs(xm - x64)) * pot0
    print(f"[gpu] max |phi_mixed - phi_fp64| = {err:.3e} V, {outer} refinement steps, {inner} FP32 CG iterations",
          flush=True)
    assert err < 1e-6
    eng.close()


def test_fp32_cycle_against_the_fp64_cycle():
    e64 = emcgpu.Engine(cfg1_deck(refine=3, electrons=60000, holes=6000), device=0, poisson_fp32=0)
    e32 = emcgpu.Engine(cfg1_deck(refine=3, electrons=60000, holes=6000), device=0, poisson_fp32=2)
    for e in (e64, e32):
        e.particles_init()
    node_q, _, _, _ = e64.deposit()
    assert np.array_equal(node_q, e32.deposit()[0])
    rhs = e64.rhs_from_charge(node_q)
    print("[gpu] FP64 CG with the FP64 and the FP32 cycle, 60x120x36", flush=True)
    x64, it64, rel64 = e64.poisson_solve(rhs, rtol=1e-10)
    x32, it32, rel32 = e32.poisson_solve(rhs, rtol=1e-10)
    assert rel64 <= 1e-10 and rel32 <= 1e-10
    # the cycle is a preconditioner: its rounding changes the search directions a
    # little, not the convergence rate
    assert abs(it32 - it64) <= 1
    lib = oracle_lib.load()
    rhs_ref = np.zeros_like(rhs)
    lib.orc_rhs(C.byref(e32.mesh), oracle_lib.i64ptr(node_q), None, oracle_lib.dptr(rhs_ref), None)
    ax = np.zeros_like(rhs)
    lib.orc_apply_full(C.byref(e32.mesh), oracle_lib.dptr(x32), oracle_lib.dptr(ax))
    assert np.linalg.norm(ax - rhs_ref) <= 1e-9 * np.linalg.norm(rhs_ref)
    # both solutions meet the 1e-10 residual: they differ by the error each may
    # have (tolerance times the condition number, about 1e3), far below a microvolt
    err = np.max(np.abs(x32 - x64)) * e32.mesh.pot0
    print(f"[gpu] max |phi_fp32cycle - phi_fp64| = {err:.3e} V, {it32} / {it64} CG iterations", flush=True)
    assert err < 1e-6
    print("[gpu] five steps with each", flush=True)
    e64.step(1)
    e32.step(1)
    assert np.max(np.abs(e32.potential() - e64.potential())) * e32.mesh.pot0 < 1e-6
    e32.step(4)  # every solve of the loop meets the tolerance, or the step fails
    for e in (e64, e32):
        e.close()


def test_step_loop_with_the_mixed_solve():
    eng = emcgpu.Engine(cfg1_deck(), device=0, poisson_fp32=True)
    ref = emcgpu.Engine(cfg1_deck(), device=0, poisson_fp32=0)
    for e in (eng, ref):
        e.particles_init()
    print("[gpu] one step with each solve", flush=True)
    eng.step(1)
    ref.step(1)
    # the first step starts both from the same population and potential: the two
    # solves of its right-hand side agree to the solver tolerance
    assert np.max(np.abs(eng.potential() - ref.potential())) * eng.mesh.pot0 < 1e-6
    eng.step(4)  # every solve of the loop meets the tolerance, or the step fails
    eng.close()
    ref.close()
