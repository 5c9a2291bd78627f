"""GPU: checkpoint files that must be refused, and a load that needs more
particle slots than the context holds.

- a truncated checkpoint and one with bytes past its end are refused before the
  context changes: the context then steps on exactly as a twin that never tried;
- a checkpoint saved under another deck (other drain voltage) or another seed is
  refused (EMC_ERR_DECK);
- a context that already sampled a small population loads a checkpoint of a
  larger one (its buffers grow) and continues exactly like the saved context."""
import os

import numpy as np
import pytest

from helpers import cfg1_deck, emcgpu
from test_gpu_parity import assert_same_population, assert_same_tallies

pytestmark = pytest.mark.gpu


def test_corrupt_checkpoints_leave_the_context_alone(tmp_path):
    path = str(tmp_path / "a.ckpt")
    src = emcgpu.Engine(cfg1_deck(), device=0)
    src.particles_init()
    src.step(2)
    src.checkpoint_save(path)
    raw = open(path, "rb").read()
    short, long_ = str(tmp_path / "short.ckpt"), str(tmp_path / "long.ckpt")
    open(short, "wb").write(raw[: len(raw) // 2])
    open(long_, "wb").write(raw + b"\0" * 16)
    a = emcgpu.Engine(cfg1_deck(), device=0)
    b = emcgpu.Engine(cfg1_deck(), device=0)
    for e in (a, b):
        e.particles_init()
        e.step(1)
    print("[gpu] loading a truncated and an overlong checkpoint", flush=True)
    for bad in (short, long_):
        with pytest.raises(emcgpu.EmcError) as ex:
            a.checkpoint_load(bad)
        assert ex.value.code == emcgpu.abi.EMC_ERR_IO
    ta, tb = a.step(2), b.step(2)
    assert_same_population(a.download(), b.download(), "after refused loads")
    assert_same_tallies(ta, tb, "after refused loads")
    np.testing.assert_array_equal(a.potential(), b.potential())
    for e in (src, a, b):
        e.close()


def test_checkpoint_of_another_deck_or_seed_is_refused(tmp_path):
    path = str(tmp_path / "a.ckpt")
    src = emcgpu.Engine(cfg1_deck(vd=0.5), device=0)
    src.particles_init()
    src.step(1)
    src.checkpoint_save(path)
    for other in (emcgpu.Engine(cfg1_deck(vd=0.4), device=0), emcgpu.Engine(cfg1_deck(vd=0.5), device=0, seed=9)):
        with pytest.raises(emcgpu.EmcError) as ex:
            other.checkpoint_load(path)
        assert ex.value.code == emcgpu.abi.EMC_ERR_DECK
        other.close()
    src.close()


def test_load_grows_the_particle_buffers(tmp_path):
    """A context that sampled a small population (ElectronNumber 4 000) loads the
    checkpoint of a large one (ElectronNumber 40 000, same mesh, doping and
    contacts): its particle buffers grow, and it continues exactly like the
    context that saved the file."""
    path = str(tmp_path / "big.ckpt")
    big = emcgpu.Engine(cfg1_deck(electrons=40000, holes=4000), device=0)
    big.particles_init()
    big.step(2)
    big.checkpoint_save(path)
    small = emcgpu.Engine(cfg1_deck(electrons=4000, holes=400), device=0, capacity_factor=1.0)
    n_small = small.particles_init()
    assert n_small < big.count()
    print("[gpu] loading a larger population into a sampled context", flush=True)
    small.checkpoint_load(path)
    # count() holds slots, dead ones included; the file holds the live particles
    assert_same_population(big.download(), small.download(), "loaded")
    tb, ts = big.step(3), small.step(3)
    assert_same_population(big.download(), small.download(), "grown")
    assert_same_tallies(tb, ts, "grown")
    big.close()
    small.close()
