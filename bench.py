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
            out[row["Counter_Name"]].setdefault(row["Kernel_Name"].split("(")[0].split("::")[-1], []).append(
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


def stream_copy_gbs(mib=1024, reps=20):
    """Measured device-to-device copy bandwidth (read + write bytes / time) of a
    plain copy kernel: what HBM gives a streaming kernel on this GPU."""
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
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * mib * 2 ** 20 / (ms * 1e-3) / 1e9


def copy_probe(mib):
    """A device-to-device copy of known size: the bytes the HBM counters should report for
    it calibrate the traffic counted for the engine's kernels."""
    import torch
    src = torch.ones(mib * 2 ** 20 // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": mib * 2 ** 20, "ok": bool(dst[-1].item() == 1.0)}))


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
    ap.add_argument("--save-checkpoint", default="", help="after the relaxation, save the state here and exit")
    ap.add_argument("--load-checkpoint", default="", help="start from this relaxed state instead of sampling")
    a = ap.parse_args()
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
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
        print(f"[bench] stream copy {extras['stream_copy_gbs']:.0f} GB/s", file=sys.stderr, flush=True)
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
