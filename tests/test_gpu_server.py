"""GPU: the serving loop (emc_server_*).  A request's drain-current windows are
the same, bit for bit, when it runs alone and when it runs in a mix with other
requests (one of its own mesh, stepped in one batch with it, and one of another
geometry, stepped at the same time on another stream); the windows equal the
currents of the same deck stepped directly through the engine, computed as
output_stat prints them; admission by device-memory footprint makes a request
wait while the budget is taken; a request stops once its current is known to the
requested precision."""
import time

import numpy as np
import pytest

from helpers import cfg1_deck, deck, emcgpu

pytestmark = pytest.mark.gpu

W, NWIN = 40, 4


def run_windows_direct(inp, seed):
    """Window currents of the deck stepped through the engine (one device): the
    drain contact's (contact 1) Current as output_stat prints it."""
    e = emcgpu.Engine(inp, device=0, seed=seed)
    e.particles_init()
    m = e.mesh
    out = []
    for _ in range(NWIN):
        t = e.step(W)
        c = emcgpu.tallies_dict(t)["contact"]
        gen = float(c[1, 1]) / emcgpu.abi.FX_CONTACT
        cat = float(c[1, 4]) / emcgpu.abi.FX_CONTACT
        out.append((gen - cat) / (m.dt * W) * m.curr0 * 1e-2)
    e.close()
    return out


def test_request_same_alone_and_in_the_mix():
    a = cfg1_deck(vd=0.5, vg=0.7, electrons=8000, holes=800)
    c = cfg1_deck(vd=0.2, vg=0.5, electrons=6000, holes=600)
    b = deck("serve_lg20", refine=1, electrons=8000, holes=800, vd=0.5, vg=0.7, lg=20, wfin=6)
    opts = dict(seed=3, window_steps=W, skip_windows=0, min_windows=NWIN, max_windows=NWIN, rel_tol=0.0)
    print("[gpu] server: request alone", flush=True)
    s = emcgpu.Server(device=0, hbm_budget_gb=16.0)
    ra = s.wait(s.submit(a, **opts), timeout=120)
    s.close()
    assert ra["state"] == emcgpu.abi.REQ_DONE and ra["windows"] == NWIN and ra["steps"] == NWIN * W
    assert not ra["converged"]
    assert ra["id_windows"] == run_windows_direct(a, 3)
    print("[gpu] server: the same request in a mix", flush=True)
    s = emcgpu.Server(device=0, hbm_budget_gb=16.0)
    ib = s.submit(b, **opts)
    ia = s.submit(a, **opts)
    ic = s.submit(c, **dict(opts, seed=5))
    rb, rm, rc = (s.wait(i, timeout=180) for i in (ib, ia, ic))
    s.close()
    for r in (rb, rm, rc):
        assert r["state"] == emcgpu.abi.REQ_DONE, r["error"]
    assert rm["id_windows"] == ra["id_windows"]
    assert rm["id_mean"] == ra["id_mean"]
    # the three ran at the same time
    assert max(r["t_start"] for r in (rb, rm, rc)) < min(r["t_done"] for r in (rb, rm, rc))
    assert rb["nodes"] != rm["nodes"]


def test_admission_by_footprint_and_precision_stop():
    a = cfg1_deck(vd=0.5, vg=0.7, electrons=8000, holes=800)
    opts = dict(window_steps=W, skip_windows=0, min_windows=NWIN, max_windows=NWIN, rel_tol=0.0)
    print("[gpu] server: a budget below one request's footprint", flush=True)
    s = emcgpu.Server(device=0, hbm_budget_gb=0.01)
    ids = [s.submit(a, seed=10 + q, **opts) for q in range(3)]
    rs = [s.wait(i, timeout=180) for i in ids]
    for r in rs:
        assert r["state"] == emcgpu.abi.REQ_DONE, r["error"]
        assert r["footprint_bytes"] > 0
    # a request always runs when nothing else does, so with a budget below one
    # footprint they run one at a time: each starts after the one before finished
    assert rs[1]["t_start"] >= rs[0]["t_done"] and rs[2]["t_start"] >= rs[1]["t_done"]
    print("[gpu] server: stop at the requested precision", flush=True)
    r = s.wait(s.submit(a, seed=4, window_steps=W, skip_windows=0, min_windows=3, max_windows=60, rel_tol=0.5),
               timeout=180)
    s.close()
    assert r["state"] == emcgpu.abi.REQ_DONE and r["converged"]
    assert 3 <= r["windows"] < 60
    assert r["id_half_width"] <= 0.5 * abs(r["id_mean"])
