#!/usr/bin/env python3
"""Benchmark: full ensemble Monte Carlo steps of the 2M-electron n-FinFET on one MI355X.

A step is the reference's density() + update_particle(): deposit (statistic,
particle_to_density, compute_cell_charge), Poisson right-hand side, PCG solve
(warm start), cell fields and surface rates, full-band flight of every particle
with GENERATE/CATCH handling.  Workload: mesh 80x160x48 cells (x, y, z) of the
n-FinFET deck, ElectronNumber=2000000, HoleNumber=500000 (plus the reference's
one carrier of each type per silicon cell), dt = 0.1 fs, T = 300 K.

The sampled population first relaxes for --relax steps (default 1000, 0.1 ps:
the carrier population and the potential reach their steady state), then W
untimed warm-up steps, then K timed steps.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build targets
one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "c-mc_electron_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

import numpy as np  # noqa: E402

import emcgpu  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per particle of one k_fly launch: read x,y,z,kx,ky,kz,E,q (64 B),
# packed cell word, itet, isym (12 B), key (8 B); write x,y,z,kx,ky,kz,E (56 B), cell word,
# itet, isym (12 B)
FLY_BYTES_PER_PARTICLE = 64 + 12 + 8 + 56 + 12
# deposit: each particle's x, y, z, charge, energy (40 B) and packed cell word (4 B) read
# once, and one read-modify-write pass over the 11 int64 node arrays (charge + 2 x 5
# window statistics: 176 B per node)
DEP_BYTES_PER_PARTICLE = 44
DEP_BYTES_PER_NODE = 176
# one CG iteration: p read, q written, x, r, z read and written, the diagonal read (FP64)
PCG_BYTES_PER_NODE_ITER = 96


def make_workload(refine, electrons, holes, tmp, qc=False):
    inp = emcgpu.write_deck(tmp, refine=refine, electrons=electrons, holes=holes, vd=0.5, vg=0.7,
                            total_step=7000, stat_step=1000, qc=int(qc), quantum_start_step=0)
    return inp


def _concat(parts):
    n = sum(q.n for q in parts)
    out = emcgpu.ParticleArrays(max(1, n))
    w = 0
    for q in parts:
        for fld in ("x", "y", "z", "kx", "ky", "kz", "energy", "charge", "i", "j", "k", "itet", "isym", "ptype",
                    "key"):
            getattr(out, fld)[w:w + q.n] = getattr(q, fld)[:q.n]
        w += q.n
    out.n = n
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on, at
    most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(eng, seconds, threads):
    """The oracle (the CPU restatement in C) runs complete steps of the same
    workload until `seconds` have passed: deposit and flight split over
    `threads` host threads by particle (each thread its own slice of the
    population, tallies and clones; the flight of one particle does not depend
    on another's, and the integer node sums add exactly), RHS, PCG (warm start
    from the GPU potential), fields and surface rates on one of them."""
    import concurrent.futures as cf
    import oracle_lib
    lib = oracle_lib.load()
    m = eng.mesh
    p = eng.download()
    live = np.nonzero(p.i[:p.n] >= 0)[0]
    p = p.subset(live)
    pot = eng.potential()
    n_ps = 0
    steps = 0
    pool = cf.ThreadPoolExecutor(max_workers=threads)
    t0 = time.perf_counter()
    while True:
        n_ps += p.n
        parts = [p.subset(ix) for ix in np.array_split(np.arange(p.n), threads)]
        dep = list(pool.map(lambda q: oracle_lib.deposit(eng, q, q.n), parts))
        node_q = np.sum([d[0] for d in dep], axis=0)
        cell_q = np.sum([d[2] for d in dep], axis=0)
        rhs = np.zeros(m.nnode)
        lib.orc_rhs(C.byref(m), oracle_lib.i64ptr(node_q), None, None, oracle_lib.dptr(rhs))
        rr = C.c_double(0)
        lib.orc_pcg(C.byref(m), oracle_lib.dptr(rhs), oracle_lib.dptr(pot), 1e-10, 100000, C.byref(rr))
        f = emcgpu.CellFieldArrays(m.ncell)
        lib.orc_cell_rho(C.byref(m), oracle_lib.i64ptr(cell_q), oracle_lib.dptr(f.rho))
        lib.orc_field(C.byref(m), oracle_lib.dptr(pot), oracle_lib.dptr(f.ex), oracle_lib.dptr(f.ey),
                      oracle_lib.dptr(f.ez))
        f.hx, f.hy, f.hz = f.ex, f.ey, f.ez
        lib.orc_surface_rates(C.byref(m), oracle_lib.dptr(f.ex), oracle_lib.dptr(f.ey), oracle_lib.dptr(f.ez),
                              oracle_lib.dptr(f.rho), oracle_lib.dptr(f.e_rough), oracle_lib.dptr(f.e_sph),
                              oracle_lib.dptr(f.h_rough), oracle_lib.dptr(f.h_sph))
        step = 10 ** 9 + steps
        outs = list(pool.map(lambda q: oracle_lib.fly_step(eng, q, q.n, f, step)[0], parts))
        p = _concat([q.subset(np.nonzero(q.i[:q.n] >= 0)[0]) for q in outs])
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    pool.shutdown()
    return n_ps / dt, dt, n_ps, steps


FLY_KERNELS = ("k_fly_fast", "k_fly_slow")


def _kernel_base_name(full):
    """'void emc::k_fly_fast<emc::FlyArgs>(emc::FlyArgs)' -> 'k_fly_fast'."""
    name = full.split("(")[0].split("<")[0].strip()
    return name.split("::")[-1].split()[-1] if name else full


def _pmc_pass(counter, args, outdir):
    """One rocprofv3 counter pass over a child run (this process has not touched
    the GPU yet).  `counter` names one counter or several separated by spaces.
    Returns {kernel name: [values per dispatch]} (one counter) or {counter: {kernel
    name: [values]}} (several), or None."""
    import csv
    import glob
    import shutil
    import subprocess
    rp = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    cmd = ["timeout", "-s", "KILL", "150", rp, "--pmc", counter, "--output-format", "csv", "-d", outdir,
           "-o", "run", "--", sys.executable, os.path.abspath(__file__)] + args
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, env=env, cwd="/tmp")
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        print(f"[bench] {counter} pass gave no counters (rc {r.returncode})", file=sys.stderr, flush=True)
        return None
    names = counter.split()
    out = {c: {} for c in names}
    for row in csv.DictReader(open(files[0])):
        if row["Counter_Name"] in out:
            out[row["Counter_Name"]].setdefault(_kernel_base_name(row["Kernel_Name"]), []).append(
                float(row["Counter_Value"]))
    return out[names[0]] if len(names) == 1 else out


# kernels of each phase of a step (the flight's headline pair first)
PHASE_KERNELS = {
    "flight": ("k_fly_fast", "k_fly_slow", "k_fly_clones_dev", "k_clone_range", "k_merge_dev"),
    "deposit": ("k_tile_hist", "k_tile_scan", "k_tile_scatter", "k_deposit_tile", "k_deposit_gather"),
    "poisson": ("k_pcg_", "k_mg_"),
}


def measure_traffic(bench_args, nsteps):
    """HBM bytes per step from FETCH_SIZE and WRITE_SIZE, each scaled by what the
    counter reports for a 512 MiB device-to-device copy (the calibration): for the
    flight's two kernels (k_fly_fast + k_fly_slow, the last step of the child run)
    and for each phase (the mean over the child's `nsteps` steps)."""
    tmp = tempfile.mkdtemp(prefix="emc_pmc_")
    mib = 512
    cf = _pmc_pass("FETCH_SIZE", ["--copy-probe", str(mib)], os.path.join(tmp, "cf"))
    cw = _pmc_pass("WRITE_SIZE", ["--copy-probe", str(mib)], os.path.join(tmp, "cw"))
    sf = _pmc_pass("FETCH_SIZE", bench_args, os.path.join(tmp, "sf"))
    sw = _pmc_pass("WRITE_SIZE", bench_args, os.path.join(tmp, "sw"))
    if not (cf and cw and sf and sw):
        return None

    def copy_kib(d):
        v = [x for k, vals in d.items() if "copyBuffer" in k for x in vals if x > 1024.0]
        return max(v) if v else None

    kib = mib * 1024.0
    f_copy, w_copy = copy_kib(cf), copy_kib(cw)
    if not f_copy or not w_copy:
        return None
    fscale, wscale = kib / f_copy, kib / w_copy

    def per_step(d):
        tot = 0.0
        for k in FLY_KERNELS:
            vals = d.get(k, [])
            if not vals:
                return None
            tot += vals[-1]  # the last (timed) step
        return tot

    def phase_bytes(phase):
        f = sum(sum(v) for k, v in sf.items() if k.startswith(PHASE_KERNELS[phase]))
        w = sum(sum(v) for k, v in sw.items() if k.startswith(PHASE_KERNELS[phase]))
        return (f * fscale + w * wscale) * 1024.0 / nsteps

    f_fly, w_fly = per_step(sf), per_step(sw)
    if f_fly is None or w_fly is None:
        return None
    out = {"bytes": (f_fly * fscale + w_fly * wscale) * 1024.0, "fetch_kib": f_fly, "write_kib": w_fly,
           "fetch_scale": fscale, "write_scale": wscale,
           "phase_bytes": {ph: phase_bytes(ph) for ph in PHASE_KERNELS}}
    # the flight is latency bound: its L2 hit rate says where the table reads are served
    hm = _pmc_pass("TCC_HIT_sum TCC_MISS_sum", bench_args, os.path.join(tmp, "hm"))
    if hm:
        hit, miss = per_step(hm["TCC_HIT_sum"]), per_step(hm["TCC_MISS_sum"])
        if hit is not None and miss is not None and hit + miss > 0:
            out["l2_hit_rate"] = hit / (hit + miss)
    return out


def stream_copy_gbs(mib=1024):
    """Measured device-to-device copy bandwidth (read + write bytes / time) of a
    plain copy: what HBM gives a streaming kernel on this GPU.  Timed in a child
    process (the engine library and torch each bring a HIP runtime; the copy uses
    torch's), see copy_probe."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-probe", str(mib)],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        print(f"[bench] stream copy child failed (rc {r.returncode}): {r.stderr[-400:]}", file=sys.stderr, flush=True)
        return None
    return json.loads(lines[-1])["gbs"]


def copy_probe(mib, reps=20):
    """A device-to-device copy of known size, repeated: the bytes the HBM counters
    report for it calibrate the traffic counted for the engine's kernels, and its
    time gives the streaming bandwidth (read + write bytes / time)."""
    import torch
    src = torch.ones(mib * 2 ** 20 // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_probe_bytes": mib * 2 ** 20, "ok": bool(dst[-1].item() == 1.0),
                      "gbs": 2.0 * mib * 2 ** 20 / (ms * 1e-3) / 1e9}))


# ---------------------------------------------------------------- bias sweep
SWEEP_VG = [0.8 * q / 7 for q in range(8)]
SWEEP_VD = [0.05 + 0.75 * q / 7 for q in range(8)]


# The sweep and serving decks: the n-FinFET of make_finfet.py with a gate
# work-function difference of +0.3 V.  With the -0.46 V of the reference's FinFET
# decks the 20x40x12-cell fin stays off up to Vg = 0.8 V (the channel sits 0.3 V
# below the source at Vg = 0.8 V); with +0.3 V its threshold falls near
# Vg = 0.3 V, so the sweep's Vg range spans off, subthreshold and on.
SWEEP_DECK = dict(phims=0.3)


def sweep_points():
    """64 bias points (Vg 0-0.8 V x Vd 0.05-0.8 V, 8 x 8) of the 20x40x12-cell
    n-FinFET (SWEEP_DECK); the electron count of a
    point follows its equilibrium inversion charge, taken linear in Vg: 50 k at
    Vg = 0 to 500 k at Vg = 0.8 V (holes a tenth; the reference's one carrier of
    each type per silicon cell on top)."""
    pts = []
    for vg in SWEEP_VG:
        for vd in SWEEP_VD:
            ne = int(50000 + 450000 * vg / 0.8)
            pts.append((vg, vd, ne))
    return pts


def sweep_engines(tmp, pts, seed0=100):
    engs = []
    for q, (vg, vd, ne) in enumerate(pts):
        d = os.path.join(tmp, f"pt{q}")
        inp = emcgpu.write_deck(d, refine=1, electrons=ne, holes=ne // 10, vd=vd, vg=vg, total_step=100000,
                                stat_step=1000, **SWEEP_DECK)
        e = emcgpu.Engine(inp, device=0, seed=seed0 + q)
        e.particles_init()
        engs.append(e)
    return engs


def window_current(t, m, w, contact=1):
    """The Current output_stat prints for `contact` over a window of w steps
    (src/mcmodel.cpp:811): contact 1 is the drain of the FinFET decks, the
    current the reference plots as its drain current (finfet/Idrain.png).  The
    mean of |contact 0| and |contact 1| (current.py) would also count the charge
    the source and drain exchange with the substrate contact, which dominates
    below threshold."""
    c = emcgpu.tallies_dict(t)["contact"]
    gen = float(c[contact, 1]) / emcgpu.abi.FX_CONTACT
    cat = float(c[contact, 4]) / emcgpu.abi.FX_CONTACT
    return (gen - cat) / (m.dt * w) * m.curr0 * 1e-2


def sweep_throughput(steps=60, solo_points=8, solo_steps=15):
    """The 64-point sweep stepped as one batch against the same points stepped one
    at a time (the plain GPU path): wall time per step of the whole sweep."""
    tmp = tempfile.mkdtemp(prefix="emc_sweep_")
    pts = sweep_points()
    t0 = time.perf_counter()
    engs = sweep_engines(tmp, pts)
    setup = time.perf_counter() - t0
    nparts = [e.count() for e in engs]
    b = emcgpu.Batch(engs)
    b.step(5)
    t0 = time.perf_counter()
    b.step(steps)
    t_batch = (time.perf_counter() - t0) / steps
    ms, _ = b.timing()
    b.close()
    # one at a time: a spread of points (every eighth), each stepped alone
    sel = list(range(0, len(engs), max(1, len(engs) // solo_points)))[:solo_points]
    per = []
    for q in sel:
        engs[q].step(3)
        t0 = time.perf_counter()
        engs[q].step(solo_steps)
        per.append((time.perf_counter() - t0) / solo_steps)
    t_solo_sweep = float(np.mean(per)) * len(engs)
    for e in engs:
        e.close()
    return {"points": len(pts), "particles_total": int(sum(nparts)), "particles_min": int(min(nparts)),
            "particles_max": int(max(nparts)), "setup_s": setup, "batched_ms_per_step": t_batch * 1e3,
            "one_at_a_time_ms_per_step": t_solo_sweep * 1e3, "speedup": t_solo_sweep / t_batch,
            "batched_particle_steps_per_s": sum(nparts) / t_batch, "batched_phase_ms": list(ms)}


def sweep_converged(window=1000, relax_steps=2000, skip=2, max_windows=30, rel=0.02, floor_frac=0.25,
                    per_stream=10, max_streams=12, solo_points=8, solo_steps=15):
    """The 64-point sweep run to a drain current known to +-rel at 95 %.

    A point's Id is the mean of the drain contact's current (window_current) over
    windows of `window` steps, each window taken after the stream's relaxation
    (`relax_steps`) and `skip` more windows.  The point is settled once the
    Student-t 95 % half-width of that mean is within rel of
    max(|Id|, floor_frac x the sweep's largest Id): a point below a quarter of the
    on-current is held to 0.5 % of it.

    The batch steps the unsettled points' streams together; a point's streams
    leave the batch once it has settled.  A point whose window spread asks for
    many more windows (an off-state point, whose current is thermal noise around
    zero) gets independent replicas: the same deck with other seeds, sampled and
    relaxed in the batch, about one stream per `per_stream` windows still needed
    (at most max_streams); its windows are pooled over its streams.  A point not
    settled after max_windows batch windows is given the windows its spread asks
    for, at the step time of the last batch.

    The plain path runs each point alone, one stream, for the same pooled number
    of windows after its relaxation and skipped windows; its time sums, per point
    and window, the window's steps times the solo step time at the point's
    particle count then (the solo step time measured on a spread of points after
    the run, fitted linear in the particle count)."""
    from scipy import stats as st
    tmp = tempfile.mkdtemp(prefix="emc_sweep_")
    pts = sweep_points()
    t_start = time.perf_counter()
    engs = sweep_engines(tmp, pts)
    m = engs[0].mesh
    nparts = [e.count() for e in engs]
    print(f"[bench] sweep: {len(engs)} points set up in {time.perf_counter() - t_start:.1f} s", file=sys.stderr,
          flush=True)
    relax_w = relax_steps // window
    # a stream: its engine, its point, and the batch window at which its windows start counting
    streams = [{"e": engs[q], "q": q, "count_from": skip + 1, "origin": True} for q in range(len(pts))]
    b = emcgpu.Batch(engs)
    t_run = time.perf_counter()
    for r0 in range(0, relax_steps, 500):
        b.step(min(500, relax_steps - r0))
        print(f"[bench] sweep relax: {min(relax_steps, r0 + 500)} of {relax_steps} steps", file=sys.stderr, flush=True)
    wins = [[] for _ in pts]
    counts = [[nparts[q], engs[q].count()] for q in range(len(pts))]
    settled_at = [None] * len(pts)
    nrep = [0] * len(pts)
    active = list(streams)
    nwin = 0
    last_ms = None
    imax = None

    def spread(q):
        x = np.array(wins[q])
        sd = x.std(ddof=1)
        scale = max(abs(x.mean()), floor_frac * imax)
        return st.t.ppf(0.975, len(x) - 1) * sd / np.sqrt(len(x)), scale, sd

    while nwin < max_windows and active:
        t0 = time.perf_counter()
        acc = b.step(window)
        last_ms = (time.perf_counter() - t0) * 1e3 / window
        nwin += 1
        for sm, t in zip(active, acc):
            if sm["origin"]:
                counts[sm["q"]].append(sm["e"].count())
            if nwin >= sm["count_from"]:
                wins[sm["q"]].append(window_current(t, m, window))
        if nwin < skip + 3:
            print(f"[bench] sweep window {nwin}", file=sys.stderr, flush=True)
            continue
        imax = max(abs(np.mean(w)) for w in wins if len(w))
        pending = sorted({sm["q"] for sm in active})
        for q in pending:
            if len(wins[q]) >= 3:
                hw, scale, _ = spread(q)
                if hw <= rel * scale:
                    settled_at[q] = nwin
        still = [sm for sm in active if settled_at[sm["q"]] is None]
        # replicas for the points whose spread asks for many more windows
        added = []
        for q in sorted({sm["q"] for sm in still}):
            if nrep[q] or len(wins[q]) < 3:
                continue
            _, scale, sd = spread(q)
            more = int(np.ceil((1.96 * sd / (rel * scale)) ** 2)) - len(wins[q])
            want = min(max_streams, int(np.ceil(more / per_stream))) - 1
            vg, vd, ne = pts[q]
            for r in range(max(0, want)):
                inp = os.path.join(tmp, f"pt{q}", "input.txt")
                e = emcgpu.Engine(inp, device=0, seed=100 + 1000 * (r + 1) + q)
                e.particles_init()
                added.append({"e": e, "q": q, "count_from": nwin + relax_w + skip + 1, "origin": False})
            nrep[q] = max(0, want)
        n_points = len({sm["q"] for sm in still})
        print(f"[bench] sweep window {nwin}: {len(pts) - n_points} of {len(pts)} points within {rel:.0%}, "
              f"{last_ms:.2f} ms per batched step of {len(active)} streams, {len(added)} replicas added",
              file=sys.stderr, flush=True)
        if still != active or added:
            b.close()
            for sm in active:
                if settled_at[sm["q"]] is not None and not sm["origin"]:
                    sm["e"].close()
            active = still + added
            if active:
                b = emcgpu.Batch([sm["e"] for sm in active])
    if active:
        b.close()
    wall_batched = time.perf_counter() - t_run
    imax = max(abs(np.mean(w)) for w in wins if len(w))
    need = []   # pooled windows each point needs
    extra_s = 0.0
    for q in range(len(pts)):
        if settled_at[q] is not None:
            need.append(len(wins[q]))
        else:
            _, scale, sd = spread(q)
            nq = max(len(wins[q]) + 1, int(np.ceil((1.96 * sd / (rel * scale)) ** 2)))
            need.append(nq)
            nstreams = 1 + nrep[q]
            extra_s = max(extra_s, (nq - len(wins[q])) / nstreams * window * last_ms * 1e-3)
    for sm in active:
        if not sm["origin"]:
            sm["e"].close()
    # one at a time: a spread of points, each stepped alone
    sel = list(range(0, len(engs), max(1, len(engs) // solo_points)))[:solo_points]
    solo = []
    for q in sel:
        engs[q].step(3)
        t0 = time.perf_counter()
        engs[q].step(solo_steps)
        solo.append((time.perf_counter() - t0) * 1e3 / solo_steps)
    fit = np.polyfit([engs[q].count() for q in sel], solo, 1)
    for e in engs:
        e.close()

    def solo_s(q):
        # relaxation at the mean of its first and last count, the skipped windows at
        # the count after relaxation, then the pooled windows one after another at
        # the counts the point's own stream had (its last count beyond them)
        c = counts[q]
        t = relax_steps * float(np.polyval(fit, 0.5 * (c[0] + c[1])))
        t += window * skip * float(np.polyval(fit, c[1]))
        run = c[2 + skip:]
        for w in range(need[q]):
            t += window * float(np.polyval(fit, run[w] if w < len(run) else c[-1]))
        return t * 1e-3

    ids = [float(np.mean(w)) for w in wins]
    t_batch_s = wall_batched + extra_s
    t_solo_s = sum(solo_s(q) for q in range(len(pts)))
    settled = sum(v is not None for v in settled_at)
    for q, (vg, vd, ne) in enumerate(pts):
        print(f"[bench] point vg={vg:.3f} vd={vd:.3f}: Id {ids[q]:.4e} A, {need[q]} windows over {1 + nrep[q]} "
              f"streams{'' if settled_at[q] is not None else ' (extrapolated)'}", file=sys.stderr, flush=True)
    return {"points": len(pts), "settled_in_run": settled, "extrapolated_points": len(pts) - settled,
            "batch_windows_run": nwin, "window_steps": window, "relax_steps": relax_steps, "skip_windows": skip,
            "replicas": int(sum(nrep)), "pooled_windows_max": int(max(need)),
            "pooled_windows_median": float(np.median(need)), "floor_frac": floor_frac,
            "batched_run_s": wall_batched, "batched_extrapolated_s": extra_s,
            "batched_sweep_s": t_batch_s, "one_at_a_time_sweep_s": t_solo_s,
            "bias_points_per_hour": len(pts) * 3600.0 / t_batch_s,
            "one_at_a_time_bias_points_per_hour": len(pts) * 3600.0 / t_solo_s,
            "speedup": t_solo_s / t_batch_s, "id_max": max(ids), "id_min": min(ids)}


# ---------------------------------------------------------------- precision
def precision_run(refine=8, electrons=8000000, holes=800000, relax_steps=100, solves=5, steps=600, window=100):
    """The 320x96x160-node-step FinFET mesh (160x320x96 cells, 4.99 M nodes) with
    8 M electrons: the FP32-bulk solve with FP64 iterative refinement and the FP64
    CG with the FP32 multigrid cycle (the default solve) against the FP64 solve
    (FP64 cycle) on the same right-hand sides (max |phi - phi_fp64| and the time
    of each), then the drain current of a run with each solve (same seed and
    initial population; windows of `window` steps, Student-t 95 % interval)."""
    from scipy import stats as st
    tmp = tempfile.mkdtemp(prefix="emc_prec_")
    inp = emcgpu.write_deck(tmp, refine=refine, electrons=electrons, holes=holes, vd=0.5, vg=0.7, total_step=100000,
                            stat_step=1000)
    t0 = time.perf_counter()
    eng = emcgpu.Engine(inp, device=0, seed=1, poisson_fp32=0)
    n0 = eng.particles_init()
    setup = time.perf_counter() - t0
    pot0 = eng.mesh.pot0
    e32 = emcgpu.Engine(inp, device=0, seed=1, poisson_fp32=2)  # solves only
    print(f"[bench] precision: {n0} particles, {eng.mesh.nnode} nodes, setup {setup:.1f} s", file=sys.stderr,
          flush=True)
    relax(eng, relax_steps, "precision relax")
    errs, t64, tmx, outer, inner, it64 = [], [], [], [], [], []
    errs32, t32, it32 = [], [], []
    for q in range(solves):
        node_q, _, _, _ = eng.deposit()
        rhs = eng.rhs_from_charge(node_q)
        x0 = eng.potential()
        t = time.perf_counter()
        x64, i64, r64 = eng.poisson_solve(rhs, x0=x0, rtol=1e-10)
        t64.append(time.perf_counter() - t)
        t = time.perf_counter()
        xm, o, i, rm = eng.poisson_solve_mixed(rhs, x0=x0, rtol=1e-10, inner_rtol=1e-5)
        tmx.append(time.perf_counter() - t)
        t = time.perf_counter()
        xc, ic, rc = e32.poisson_solve(rhs, x0=x0, rtol=1e-10)
        t32.append(time.perf_counter() - t)
        errs.append(float(np.max(np.abs(xm - x64)) * pot0))
        errs32.append(float(np.max(np.abs(xc - x64)) * pot0))
        outer.append(o)
        inner.append(i)
        it64.append(i64)
        it32.append(ic)
        eng.step(20)
        print(f"[bench] precision solve {q}: fp64 {t64[-1]*1e3:.1f} ms ({i64} it); mixed max err {errs[-1]:.3e} V, "
              f"{tmx[-1]*1e3:.1f} ms ({o} outer, {i} FP32 it); fp32 cycle max err {errs32[-1]:.3e} V, "
              f"{t32[-1]*1e3:.1f} ms ({ic} it)", file=sys.stderr, flush=True)
    eng.close()
    e32.close()

    def id_run(fp32):
        e = emcgpu.Engine(inp, device=0, seed=1, poisson_fp32=fp32)
        e.particles_init()
        m = e.mesh
        relax(e, relax_steps, "precision id relax")
        w = []
        t = time.perf_counter()
        for k in range(steps // window):
            w.append(window_current(e.step(window), m, window))
        el = time.perf_counter() - t
        ms, it = e.timing()
        e.close()
        x = np.array(w)
        hw = st.t.ppf(0.975, len(x) - 1) * x.std(ddof=1) / np.sqrt(len(x))
        return {"id_mean": float(x.mean()), "id_half_width": float(hw), "windows": len(x),
                "ms_per_step": el / steps * 1e3, "poisson_ms": float(ms[1])}

    r64 = id_run(0)
    rmx = id_run(1)
    r32 = id_run(2)
    print(f"[bench] precision id: fp64 {r64}, mixed {rmx}, fp32 cycle {r32}", file=sys.stderr, flush=True)
    return {"particles": int(n0), "setup_s": setup,
            "max_abs_err_v": max(errs), "errs_v": errs, "fp64_solve_ms": float(np.mean(t64)) * 1e3,
            "mixed_solve_ms": float(np.mean(tmx)) * 1e3, "fp64_cg_iters": it64, "refine_steps": outer,
            "fp32_cg_iters": inner, "id_fp64": r64, "id_mixed": rmx,
            "id_diff_in_combined_half_widths": abs(r64["id_mean"] - rmx["id_mean"]) /
            max(1e-300, np.hypot(r64["id_half_width"], rmx["id_half_width"])),
            "fp32_cycle": {"max_abs_err_v": max(errs32), "errs_v": errs32, "solve_ms": float(np.mean(t32)) * 1e3,
                           "cg_iters": it32, "id": r32,
                           "id_diff_in_combined_half_widths": abs(r64["id_mean"] - r32["id_mean"]) /
                           max(1e-300, np.hypot(r64["id_half_width"], r32["id_half_width"]))}}


# ---------------------------------------------------------------- serving
def serve_stream(nreq=12, rate=2.0, budget_gb=24.0, window=500, seed=7, abs_tol=2e-9):
    """A Poisson stream (mean `rate` requests/s) of bias/geometry requests: gate
    length 10-30 nm, fin width 4-8 nm, 2 or 3 grid lines per nm (meshes of 45 k
    to 420 k nodes), Vg 0.3-0.8 V, Vd 0.1-0.8 V, 20 k-100 k electrons, the
    SWEEP_DECK device; each runs until its drain current is
    within +-2 % at 95 %, or within abs_tol (about 1 % of the on-state current)
    for requests below threshold (windows of `window` steps, the first two dropped),
    or 40 windows.  Latency = submission to result."""
    rng = np.random.default_rng(seed)
    tmp = tempfile.mkdtemp(prefix="emc_serve_")
    srv = emcgpu.Server(device=0, hbm_budget_gb=budget_gb)
    ids = []
    t0 = time.perf_counter()
    next_t = 0.0
    for q in range(nreq):
        next_t += rng.exponential(1.0 / rate)
        lg, wf, ref = int(rng.integers(10, 31)), int(rng.integers(4, 9)), int(rng.integers(2, 4))
        vg, vd, ne = float(rng.uniform(0.3, 0.8)), float(rng.uniform(0.1, 0.8)), int(rng.integers(20000, 100001))
        inp = emcgpu.write_deck(os.path.join(tmp, f"r{q}"), refine=ref, electrons=ne, holes=ne // 10, vd=vd, vg=vg,
                                lg=lg, wfin=wf, total_step=100000, stat_step=1000, **SWEEP_DECK)
        while time.perf_counter() - t0 < next_t:
            time.sleep(0.005)
        ids.append(srv.submit(inp, seed=1000 + q, window_steps=window, skip_windows=2, min_windows=3,
                              max_windows=40, rel_tol=0.02, abs_tol=abs_tol))
    res = [None] * nreq
    seen = [0] * nreq
    while any(r is None for r in res):
        for q, i in enumerate(ids):
            if res[q] is None:
                r = srv.poll(i)
                if r["windows"] > seen[q]:
                    seen[q] = r["windows"]
                    print(f"[bench] serve request {q}: window {r['windows']}", file=sys.stderr, flush=True)
                if r["state"] in (emcgpu.abi.REQ_DONE, emcgpu.abi.REQ_FAILED):
                    res[q] = r
                    print(f"[bench] serve request {q}: {r['windows']} windows, Id {r['id_mean']:.4e} A "
                          f"+- {r['id_half_width']:.2e}, {r['t_done'] - r['t_submit']:.1f} s", file=sys.stderr,
                          flush=True)
        time.sleep(0.2)
    wall = time.perf_counter() - t0
    srv.close()
    ok = [r for r in res if r["state"] == emcgpu.abi.REQ_DONE]
    lat = np.array([r["t_done"] - r["t_submit"] for r in ok])
    wait = np.array([r["t_start"] - r["t_submit"] for r in ok])
    return {"requests": nreq, "done": len(ok), "converged": int(sum(r["converged"] for r in ok)),
            "arrival_rate_per_s": rate, "hbm_budget_gb": budget_gb,
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "admission_wait_p50_s": float(np.percentile(wait, 50)) if len(wait) else None,
            "requests_per_hour": len(ok) * 3600.0 / wall, "wall_s": wall,
            "nodes_min": int(min(r["nodes"] for r in ok)) if ok else None,
            "nodes_max": int(max(r["nodes"] for r in ok)) if ok else None,
            "particles_max": int(max(r["particles"] for r in ok)) if ok else None,
            "errors": [r["error"] for r in res if r["state"] != emcgpu.abi.REQ_DONE]}


def relax(eng, nsteps, tag="relax"):
    done = 0
    while done < nsteps:
        k = min(250, nsteps - done)
        eng.step(k)
        done += k
        print(f"[bench] {tag}: {done} of {nsteps} steps", file=sys.stderr, flush=True)


def roofline(kernel, nbytes, ms, per_unit, traffic=None):
    if not ms or ms <= 0:
        return None
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": nbytes, "per_unit": per_unit}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--relax", type=int, default=1000,
                    help="steps the sampled population flies before the warm-up (steady state)")
    ap.add_argument("--refine", type=int, default=4)
    ap.add_argument("--electrons", type=int, default=2000000)
    ap.add_argument("--holes", type=int, default=500000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the cores this process may use, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes (roofline.traffic = null)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the MultipleRefresh, heat-phase and stream-copy measurements")
    ap.add_argument("--qc", action="store_true",
                    help="quantum correction on (Flag_QuantumCorrection=1 from step 0; not the default workload)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--sweep", action="store_true",
                    help="only the 64-point bias sweep run to +-2 %% Id (bias-points/hour), one JSON line")
    ap.add_argument("--serve", action="store_true",
                    help="only the serving stream (p50/p99 latency, requests/hour), one JSON line")
    ap.add_argument("--precision", action="store_true",
                    help="only the mixed-precision Poisson workload (4.99 M nodes, 8 M electrons), one JSON line")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep throughput measurement")
    ap.add_argument("--save-checkpoint", default="", help="after the relaxation, save the state here and exit")
    ap.add_argument("--load-checkpoint", default="", help="start from this relaxed state instead of sampling")
    a = ap.parse_args()
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if a.sweep:
        r = sweep_converged()
        print(json.dumps({"metric": "bias-points/hour at +-2% Id (95% CI)", "value": r["bias_points_per_hour"],
                          "unit": "bias-points/hour", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": "64-point sweep of the 20x40x12-cell n-FinFET, "
                                                          "one batch", **r}}))
        return
    if a.precision:
        r = precision_run()
        print(json.dumps({"metric": "max |phi_mixed - phi_fp64|", "value": r["max_abs_err_v"], "unit": "V",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f32 bulk + f64 refinement",
                          "data": "synthetic", "config": {"workload": "n-FinFET 160x320x96 cells (4.99 M nodes), "
                                                          "ElectronNumber=8000000", **r}}))
        return
    if a.serve:
        r = serve_stream()
        print(json.dumps({"metric": "p50/p99 latency", "value": r["latency_p50_s"], "unit": "s", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": "Poisson stream of FinFET bias/geometry requests", **r}}))
        return
    wl = ["--refine", str(a.refine), "--electrons", str(a.electrons), "--holes", str(a.holes)] + \
        (["--qc"] if a.qc else [])
    traffic = None
    ckpt = a.load_checkpoint
    if not a.no_traffic and not ckpt and not a.save_checkpoint:
        # before this process touches the GPU: a child relaxes the population and
        # saves it; the counter passes (children under rocprofv3) and this process
        # start from that state
        import subprocess
        ckpt = os.path.join(tempfile.mkdtemp(prefix="emc_relaxed_"), "relaxed.ckpt")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--save-checkpoint", ckpt,
                            "--relax", str(a.relax)] + wl)
        if r.returncode != 0:
            sys.exit(r.returncode)
        traffic = measure_traffic(["--load-checkpoint", ckpt, "--steps", "1", "--warmup", "1", "--no-cpu",
                                   "--no-traffic", "--no-extras"] + wl, 2)
        print(f"[bench] traffic {traffic}", file=sys.stderr, flush=True)
    tmp = tempfile.mkdtemp(prefix="emc_bench_")
    inp = make_workload(a.refine, a.electrons, a.holes, tmp, qc=a.qc)
    print(f"[bench] deck {inp}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    eng = emcgpu.Engine(inp, device=0, seed=1)
    if ckpt:
        eng.checkpoint_load(ckpt)
        n0 = eng.count()
        relaxed = a.relax
    else:
        n0 = eng.particles_init()
        relaxed = 0
    print(f"[bench] {n0} particles on {eng.mesh.cnx}x{eng.mesh.cny}x{eng.mesh.cnz} cells, "
          f"{eng.mesh.nnode} nodes, setup {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    if not ckpt:
        relax(eng, a.relax)
        relaxed = a.relax
        if a.save_checkpoint:
            eng.checkpoint_save(a.save_checkpoint)
            print(f"[bench] relaxed state saved to {a.save_checkpoint}", file=sys.stderr, flush=True)
            eng.close()
            return
    for w in range(a.warmup):
        eng.step(1)
    print(f"[bench] {a.warmup} warm-up steps done", file=sys.stderr, flush=True)
    n_start = eng.count()
    phase = np.zeros(5)
    pcg_it = []
    n_ps = 0
    full_loop = events = 0
    t0 = time.perf_counter()
    for s in range(a.steps):
        n_ps += eng.count()
        t = eng.step(1)
        ms, it = eng.timing()
        phase += ms
        pcg_it.append(it)
        full_loop += t.n_full_loop
        events += t.n_events
    t1 = time.perf_counter()
    elapsed = t1 - t0
    phase /= max(1, a.steps)
    value = n_ps / elapsed
    n_avg = n_ps / max(1, a.steps)
    nodes = eng.mesh.nnode
    it_mean = float(np.mean(pcg_it))
    print(f"[bench] {a.steps} steps in {elapsed:.3f}s; per step ms: deposit+rhs {phase[0]:.3f} "
          f"pcg {phase[1]:.3f} field {phase[2]:.3f} fly(all) {phase[3]:.3f} k_fly {phase[4]:.3f}; "
          f"pcg iters {it_mean:.1f}", file=sys.stderr, flush=True)
    extras = {}
    if not a.no_extras:
        # MultipleRefresh of the same population, once (the bench deck, like the
        # reference's plain configuration, runs without it), then the heating phase
        t_mr = time.perf_counter()
        mr_before, mr_after = eng.multiple_refresh(10 ** 6)
        extras["multiple_refresh_ms"] = (time.perf_counter() - t_mr) * 1e3
        print(f"[bench] multiple refresh {mr_before} -> {mr_after} particles in "
              f"{extras['multiple_refresh_ms']:.2f} ms", file=sys.stderr, flush=True)
        eng.heat_begin()
        eng.heat_step(2)
        nh = 20
        t_h = time.perf_counter()
        eng.heat_step(nh)
        extras["heat_phase_ms_per_step"] = (time.perf_counter() - t_h) * 1e3 / nh
        print(f"[bench] heating phase {extras['heat_phase_ms_per_step']:.3f} ms per step", file=sys.stderr,
              flush=True)
        extras["stream_copy_gbs"] = stream_copy_gbs()
        if extras["stream_copy_gbs"]:
            print(f"[bench] stream copy {extras['stream_copy_gbs']:.0f} GB/s", file=sys.stderr, flush=True)
    if not a.no_extras and not a.no_sweep:
        try:
            extras["sweep"] = sweep_throughput()
            print(f"[bench] sweep {extras['sweep']}", file=sys.stderr, flush=True)
        except Exception as ex:  # reported, not fatal: the headline is the single-device step
            extras["sweep"] = {"error": str(ex)}
    cpu = None
    if not a.no_cpu and a.cpu_seconds > 0:
        thr = a.cpu_threads or cpu_threads()
        v, dt, nps, st = cpu_baseline(eng, a.cpu_seconds, thr)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": thr, "kind": "port",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "sample": f"{st} complete steps of the same relaxed population ({nps} particle-steps, {dt:.1f} s) "
                         f"by the oracle in C, deposit and flight on {thr} threads split by particle, "
                         f"PCG and fields on one"}
        print(f"[bench] cpu baseline {v:.4g} particle-steps/s over {dt:.1f}s on {thr} threads", file=sys.stderr,
              flush=True)
    pb = traffic.get("phase_bytes", {}) if traffic else {}
    fly_bytes = FLY_BYTES_PER_PARTICLE * n_avg
    fly = roofline("k_fly_fast + k_fly_slow", fly_bytes, phase[4], f"{FLY_BYTES_PER_PARTICLE} B per particle",
                   traffic["bytes"] if traffic else None)
    if fly:
        fly["traffic_counters"] = traffic
        fly["l2_hit_rate"] = traffic.get("l2_hit_rate") if traffic else None
        fly["bytes_per_particle"] = FLY_BYTES_PER_PARTICLE
    dep_bytes = DEP_BYTES_PER_PARTICLE * n_avg + DEP_BYTES_PER_NODE * nodes
    pcg_bytes = PCG_BYTES_PER_NODE_ITER * nodes * it_mean
    phases = {
        "deposit": roofline("tile sort + k_deposit_tile + k_deposit_gather (deposit + rhs phase)", dep_bytes,
                            phase[0], f"{DEP_BYTES_PER_PARTICLE} B per particle + {DEP_BYTES_PER_NODE} B per node",
                            pb.get("deposit")),
        "poisson": roofline("MG-PCG (k_pcg_*, k_mg_*)", pcg_bytes, phase[1],
                            f"{PCG_BYTES_PER_NODE_ITER} B per node per CG iteration", pb.get("poisson")),
        "flight_all": roofline("flight incl. clone passes and merge", fly_bytes, phase[3],
                               f"{FLY_BYTES_PER_PARTICLE} B per particle", pb.get("flight")),
    }
    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"n-FinFET {eng.mesh.cnx}x{eng.mesh.cny}x{eng.mesh.cnz} cells, "
                               f"{n_start} particles after {relaxed} relaxation steps (ElectronNumber={a.electrons}, "
                               f"HoleNumber={a.holes}), dt 0.1 fs, full step: deposit + PCG Poisson + fields + flight"
                               + (" + quantum-correction potential" if a.qc else ""),
                   "quantum_correction": bool(a.qc), "relax_steps": relaxed,
                   "poisson_solve": "FP64 CG to ||r|| <= 1e-10 ||b||, preconditioned by a multigrid cycle in FP32",
                   "particles": int(n_start), "nodes": int(nodes),
                   "phase_ms": {"deposit_rhs": phase[0], "poisson": phase[1], "field": phase[2],
                                "flight_total": phase[3], "k_fly": phase[4]},
                   "pcg_iters_mean": it_mean,
                   "full_event_loop_fraction": full_loop / max(1, n_ps),
                   "events_per_particle_step": events / max(1, n_ps), **extras},
        "roofline": fly,
        "phase_rooflines": phases,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
