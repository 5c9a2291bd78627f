"""emc_poisson_solve_noalloc (work vectors in the library's static device
array) against emc_poisson_solve on the same inputs: identical solution,
iteration count and residual, on two calls in a row, the second on a smaller
mesh (another context on the same GPU, so it finds the first call's contents in
the shared array), with the multigrid cycle in FP32 (the default) and in FP64."""
import numpy as np
import pytest

from helpers import deck, emcgpu


def _case(refine, electrons, holes, seed, fp32_cycle):
    eng = emcgpu.Engine(deck("noalloc", refine=refine, electrons=electrons, holes=holes, vd=0.5, vg=0.7), device=0,
                        poisson_fp32=2 if fp32_cycle else 0)
    eng.particles_init()
    node_q, _, _, _ = eng.deposit()
    rhs = eng.rhs_from_charge(node_q)
    rng = np.random.default_rng(seed)
    x0 = rng.normal(0.0, 1e-3, eng.mesh.nnode)
    return eng, rhs, x0


@pytest.mark.gpu
@pytest.mark.parametrize("fp32_cycle", [True, False])
def test_noalloc_solve_equals_the_allocating_solve(fp32_cycle):
    for refine, ne, nh, seed in ((2, 40000, 4000, 1), (1, 10000, 1000, 2)):
        eng, rhs, x0 = _case(refine, ne, nh, seed, fp32_cycle)
        print(f"[gpu] noalloc solve, refine {refine}, {eng.mesh.nnode} nodes, FP32 cycle {fp32_cycle}", flush=True)
        xa, ia, ra = eng.poisson_solve(rhs, x0=x0, rtol=1e-10)
        xb, ib, rb = eng.poisson_solve_noalloc(rhs, x0=x0, rtol=1e-10)
        assert ia == ib and ia > 0
        assert ra == rb and rb <= 1e-10
        np.testing.assert_array_equal(xa, xb)
        eng.close()


@pytest.mark.gpu
def test_noalloc_solve_refuses_a_mesh_beyond_the_array():
    eng = emcgpu.Engine(deck("noalloc_big", refine=5, electrons=20000, holes=2000, vd=0.5, vg=0.7), device=0)
    assert eng.mesh.nnode > 1_000_000
    rhs = np.zeros(eng.mesh.nnode)
    with pytest.raises(emcgpu.EmcError) as e:
        eng.poisson_solve_noalloc(rhs)
    assert e.value.code == emcgpu.abi.EMC_ERR_CAPACITY
    eng.close()
