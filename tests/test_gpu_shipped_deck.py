"""GPU, whole run: the reference's own FinFET deck (finfet/input.txt, ldg.txt,
lgrid.txt, copied to c-mc_electron_amd/decks/finfet_shipped) run for its 7 000
steps of 0.2 fs with MultipleRefresh every 50 steps.  The window-mean potential of
steps 6 000..6 999 (pot6999, output_stat) equals the one in the result file the
reference ships for its own run of this deck (finfet/pot6999.vtr, values printed
to 6 significant digits; tests/golden/finfet_pot6999.npz) within 20 uV at every
node.  Measured: 0.8 uV rms, 5.1 uV at most, the print precision of the file.
The potential of this deck is set by its doping, oxide and contacts: the first
solve from the initial (neutral) population already agrees within 11 uV.  The
check therefore covers the deck, mesh, operator, deposit, solve and window
average against the reference's own output, not the carrier transport."""
import os

import numpy as np
import pytest

from helpers import GOLDEN_DIR, PKG, emcgpu

pytestmark = pytest.mark.gpu


def test_shipped_deck_potential_matches_the_reference_run():
    ref = np.load(os.path.join(GOLDEN_DIR, "finfet_pot6999.npz"))["potential"]  # [k][j][i]
    e = emcgpu.Engine(os.path.join(PKG, "decks", "finfet_shipped", "input.txt"), device=0, seed=1)
    try:
        m = e.mesh
        assert (m.pnz, m.pny, m.pnx) == ref.shape
        e.particles_init()
        stat_step = 1000
        for w in range(7):  # windows of stat_step steps, as run() resets them (set_stat_zero)
            print(f"[gpu] shipped deck, steps {w * stat_step}..{w * stat_step + stat_step - 1}", flush=True)
            e.tallies_reset()
            e.step(stat_step)
        win = e.stat_window()
        assert win["steps"] == stat_step
        got = (win["stat_pot"] / stat_step * m.pot0).reshape(m.pny, m.pnx, m.pnz).transpose(2, 0, 1)
        d = got - ref
        rms, worst = float(np.sqrt((d * d).mean())), float(np.abs(d).max())
        print(f"[gpu] pot6999 against the reference: rms {rms * 1e6:.2f} uV, max {worst * 1e6:.2f} uV", flush=True)
        assert worst < 20e-6 and rms < 5e-6
    finally:
        e.close()
