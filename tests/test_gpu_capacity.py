"""GPU: a population that outgrows its particle slots keeps every particle.
Below threshold the FinFET's contacts generate more carriers than they catch
and, without MultipleRefresh, the population grows by about 5 % per 1 000
steps.  A context sampled with no spare slots (capacity_factor 1) moves its
particles into larger buffers between steps, alone and in a batch, and steps on
exactly like a context with room to spare: the same potential and population,
bit for bit, and no particle dropped for capacity (drop reason 8)."""
import numpy as np
import pytest

from helpers import deck, emcgpu
from test_gpu_parity import assert_same_population

pytestmark = pytest.mark.gpu

STEPS = 3000


def grow_deck():
    return deck("grow", refine=1, electrons=20000, holes=2000, vd=0.05, vg=0.0, phims=0.3)


def grew(n0, p):
    """Whether the live population passed the point where a context sampled with
    capacity_factor 1 (n0 + 1024 slots) must move into larger buffers: within one
    clone buffer (max(4096, slots / 16)) of its slots."""
    cap0 = n0 + 1024
    return int((p.i[:p.n] >= 0).sum()) > cap0 - max(4096, cap0 // 16)


def test_population_outgrows_its_slots():
    tight = emcgpu.Engine(grow_deck(), device=0, seed=3, capacity_factor=1.0)
    roomy = emcgpu.Engine(grow_deck(), device=0, seed=3, capacity_factor=4.0)
    n0 = tight.particles_init()
    assert roomy.particles_init() == n0
    print(f"[gpu] capacity: {STEPS} steps from {n0} particles, tight and roomy", flush=True)
    drops = 0
    for _ in range(STEPS // 100):
        tt, tr = tight.step(100), roomy.step(100)
        drops += int(tt.drops[8]) + int(tr.drops[8])
    assert drops == 0
    pt, pr = tight.download(), roomy.download()
    assert grew(n0, pt), "the population did not outgrow the tight context's first slots"
    assert_same_population(pt, pr, "grown")
    np.testing.assert_array_equal(tight.potential(), roomy.potential())
    tight.close()
    roomy.close()


def test_population_outgrows_its_slots_in_a_batch():
    a = emcgpu.Engine(grow_deck(), device=0, seed=5, capacity_factor=1.0)
    b = emcgpu.Engine(deck("grow_on", refine=1, electrons=20000, holes=2000, vd=0.5, vg=0.8, phims=0.3), device=0,
                      seed=6, capacity_factor=1.0)
    solo = emcgpu.Engine(grow_deck(), device=0, seed=5, capacity_factor=4.0)
    n0 = a.particles_init()
    b.particles_init()
    solo.particles_init()
    print(f"[gpu] capacity: {STEPS} batched steps", flush=True)
    bat = emcgpu.Batch([a, b])
    drops = 0
    for _ in range(STEPS // 100):
        for t in bat.step(100):
            drops += int(t.drops[8])
        drops += int(solo.step(100).drops[8])
    bat.close()
    assert drops == 0
    pa = a.download()
    assert grew(n0, pa)
    assert_same_population(pa, solo.download(), "batched, grown")
    np.testing.assert_array_equal(a.potential(), solo.potential())
    for e in (a, b, solo):
        e.close()
