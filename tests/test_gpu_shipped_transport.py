"""GPU, whole run: the reference's own FinFET deck (finfet/input.txt, copied to
c-mc_electron_amd/decks/finfet_shipped) run end to end by `emc_run input.txt`:
7 000 transient steps of 0.2 fs from the sampled population, MultipleRefresh every
50 steps, then the 7 000-step heating phase at the window-mean potential.  Its
contact currents and heat profile are set against the reference's own run of the
same deck, read off the result files it ships (tests/golden/finfet_fields6999.npz,
written by tests/golden/make_shipped_fields.py from finfet/ids_vs_step.png,
finfet/ids_vs_step2.png and finfet/heat6999.vtr):

                                   reference   here (seed 1)   bound
  contact 1, window 0..999         7.788 pA    7.768 pA        1 %
  contact 1, window 1000..1999     3.730 pA    3.820 pA        4 %
  contact 1, steady (12 windows)   3.427 pA    3.520 pA        4.5 %
  contact 0, window 0..999        -3.443 pA   -3.420 pA        2 %
  contact 0, windows 3..7         -3.463 pA   -3.581 pA        4.5 %
  heat, y profile correlation      1           0.999           > 0.99
  heat, total                      1           1.08            15 %

Seeds 1 and 2, runs without MultipleRefresh and the all-FP64 solve give the same
picture: the first window agrees within 0.3 % (contact 1), the steady drain
current is 2.5 to 2.7 % above the reference's and the source current 3 to 3.4 %
above.  The first window is the
discharge of the sampled population into the contacts and is set by the initial
state, the fields and the flight; the steady windows add the contacts' particle
exchange and MultipleRefresh over thousands of steps.  The bounds hold the
measured offsets with room for the run-to-run spread (0.5 % per window); a
change of the transport physics by a few percent, or of the contact or
MultipleRefresh bookkeeping, fails them.  Steady contact 1 uses the reference's
two markers at each of the steps 3000..7000 (transient and heating phase
overlap there) and its heating-phase markers at 1000 and 2000."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN_DIR, PKG, REPO, emcgpu

pytestmark = pytest.mark.gpu


def _grid(path):
    vals = [float(v) for v in open(path).read().split()]
    out, p = [], 0
    for _ in range(3):
        n = int(vals[p])
        out.append(np.array(vals[p + 1:p + 1 + n]))
        p += 1 + n
    return out


def _y_profile(nodes, x, y, z):
    """Volume-weighted mean over x and z of the cell averages of a node field [i][j][k]."""
    c = sum(nodes[a:nodes.shape[0] - 1 + a, b:nodes.shape[1] - 1 + b, d:nodes.shape[2] - 1 + d]
            for a in (0, 1) for b in (0, 1) for d in (0, 1)) / 8
    vol = np.einsum("i,j,k->ijk", np.diff(x), np.diff(y), np.diff(z))
    return (c * vol).sum(axis=(0, 2)) / vol.sum(axis=(0, 2))


def test_shipped_deck_currents_and_heat_against_the_reference_run(tmp_path):
    G = np.load(os.path.join(GOLDEN_DIR, "finfet_fields6999.npz"))
    deck = os.path.join(PKG, "decks", "finfet_shipped")
    for f in ("input.txt", "ldg.txt", "lgrid.txt"):
        shutil.copy(os.path.join(deck, f), tmp_path)
    exe = os.path.join(REPO, "c-mc_electron_amd", "bin", "emc_run")
    print("[gpu] shipped deck: emc_run, 7000 transient + 7000 heating-phase steps", flush=True)
    r = subprocess.run([exe, "input.txt", "--seed", "1"], cwd=tmp_path, stdout=subprocess.DEVNULL,
                       stderr=subprocess.PIPE, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    win = emcgpu.read_current(str(tmp_path / "data" / "current"))
    assert [w["step"] for w in win] == [999 + 1000 * q for q in range(7)] * 2
    c1 = np.array([w["current"][1] for w in win])
    c0 = np.array([w["current"][0] for w in win])
    tr1, heat1, tr0 = c1[:7], c1[7:], c0[:7]
    ref_both = G["ids_vs_step_contact1"]  # [step][upper, lower]
    ref0 = G["current_contact0"]
    ref_steady = np.concatenate([ref_both[2:].ravel(), ref_both[:2, 1]])
    ours_steady = np.concatenate([tr1[2:], heat1])
    checks = {
        "contact 1 window 0": (tr1[0] / ref_both[0, 0] - 1, 0.01),
        "contact 1 window 1": (tr1[1] / ref_both[1, 0] - 1, 0.04),
        "contact 1 steady": (ours_steady.mean() / ref_steady.mean() - 1, 0.045),
        "contact 0 window 0": (tr0[0] / ref0[0] - 1, 0.02),
        "contact 0 windows 3..7": (tr0[2:].mean() / ref0[2:].mean() - 1, 0.045),
    }
    for k, (d, _) in checks.items():
        print(f"[gpu] {k}: {100 * d:+.2f} % against the reference run", flush=True)
    # heating phase: the heat6999 node field (electron + hole columns) and its y profile
    x, y, z = (v * 1e-6 for v in _grid(str(tmp_path / "lgrid.txt")))
    nx, ny, nz = len(x), len(y), len(z)
    raw = np.loadtxt(tmp_path / "data" / "heat6999")
    heat = np.zeros((nx, ny, nz))
    idx = (raw[:, 0].astype(int), raw[:, 1].astype(int), raw[:, 2].astype(int))
    for col in (4, 6):
        np.add.at(heat, idx, -raw[:, col])
    ref_heat = G["heat"].astype(float)
    corr = float(np.corrcoef(_y_profile(heat, x, y, z), _y_profile(ref_heat, x, y, z))[0, 1])
    total = float(heat.sum() / ref_heat.sum())
    print(f"[gpu] heat: y-profile correlation {corr:.4f}, total {total:.3f} x the reference", flush=True)
    for k, (d, tol) in checks.items():
        assert abs(d) < tol, f"{k}: {100 * d:+.2f} % (bound {100 * tol:.1f} %)"
    assert corr > 0.99
    assert abs(total - 1) < 0.15
