"""GPU: a batch of devices that share the mesh (a bias sweep) stepped together,
every kernel one launch for all of them (emc_batch_*), against each device
stepped alone with emc_step: after the same steps, every device's potential,
population (sorted by key), tallies and statistics window are identical bit for
bit.  The devices differ in drain and gate bias and in particle count (uneven
populations, as the equilibrium charge of each bias point sets them).  Both
precisions of the multigrid cycle (FP32, the default, and FP64) are run."""
import numpy as np
import pytest

from helpers import cfg1_deck, emcgpu, sort_by_key

pytestmark = pytest.mark.gpu

POINTS = [(0.5, 0.7, 6000), (0.05, 0.0, 9000), (0.8, 0.4, 3000), (0.3, 0.8, 12000)]
STEPS = 6


def engines(mode):
    out = []
    for vd, vg, ne in POINTS:
        e = emcgpu.Engine(cfg1_deck(vd=vd, vg=vg, electrons=ne, holes=ne // 10), device=0, seed=7, poisson_fp32=mode)
        e.particles_init()
        out.append(e)
    return out


def snapshot(e):
    p = sort_by_key(e.download())
    w = e.stat_window()
    return {"pot": e.potential(), "p": p, "win": w}


@pytest.mark.parametrize("mode", [2, 0])
def test_batch_equals_each_device_alone(mode):
    print(f"[gpu] batch of 4 bias points, poisson_fp32 {mode}", flush=True)
    bat_eng = engines(mode)
    b = emcgpu.Batch(bat_eng)
    acc_b = [b.step(2) for _ in range(STEPS // 2)]
    ms, imax = b.timing()
    assert imax > 0 and ms[4] > 0
    got = [snapshot(e) for e in bat_eng]
    b.close()
    print("[gpu] the same points one at a time", flush=True)
    solo_eng = engines(mode)
    acc_s = [[e.step(2) for _ in range(STEPS // 2)] for e in solo_eng]
    want = [snapshot(e) for e in solo_eng]
    for q in range(len(POINTS)):
        g, w = got[q], want[q]
        assert g["p"].n == w["p"].n, q
        for fld in ("x", "y", "z", "kx", "ky", "kz", "energy", "charge", "i", "j", "k", "itet", "isym", "ptype", "key"):
            np.testing.assert_array_equal(getattr(g["p"], fld)[:g["p"].n], getattr(w["p"], fld)[:w["p"].n], err_msg=f"{q} {fld}")
        np.testing.assert_array_equal(g["pot"], w["pot"])
        assert g["win"]["steps"] == w["win"]["steps"] == STEPS
        for k in ("stat_pot", "node_stat", "cell_count", "rcur", "lcur"):
            np.testing.assert_array_equal(np.asarray(g["win"][k]), np.asarray(w["win"][k]), err_msg=f"{q} {k}")
        for s in range(STEPS // 2):
            tb, ts = emcgpu.tallies_dict(acc_b[s][q]), emcgpu.tallies_dict(acc_s[q][s])
            for k in tb:
                np.testing.assert_array_equal(np.asarray(tb[k]), np.asarray(ts[k]), err_msg=f"{q} step {s} {k}")
    # the points are different devices: their potentials and currents differ
    assert not np.array_equal(got[0]["pot"], got[1]["pot"])
    for e in bat_eng + solo_eng:
        e.close()


def test_batch_refuses_mixed_meshes():
    a = emcgpu.Engine(cfg1_deck(), device=0)
    c = emcgpu.Engine(cfg1_deck(refine=2, electrons=4000, holes=400), device=0)
    a.particles_init()
    c.particles_init()
    with pytest.raises(emcgpu.EmcError) as e:
        emcgpu.Batch([a, c])
    assert e.value.code == emcgpu.abi.EMC_ERR_ARG
    c.close()
    # one FP64 CG for all: the cycle's precision is shared, the mixed solve not batched
    for mode in (0, 1):
        d = emcgpu.Engine(cfg1_deck(), device=0, poisson_fp32=mode)
        d.particles_init()
        with pytest.raises(emcgpu.EmcError) as e:
            emcgpu.Batch([a, d])
        assert e.value.code == emcgpu.abi.EMC_ERR_ARG
        d.close()
    a.close()
