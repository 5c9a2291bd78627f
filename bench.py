#!/usr/bin/env python3
"""Benchmark: full ensemble Monte Carlo steps of the 2M-electron n-FinFET on one MI355X.

A step is the reference's density() + update_particle(): deposit (statistic,
particle_to_density, compute_cell_charge), Poisson right-hand side, PCG solve
(warm start), cell fields and surface rates, full-band flight of every particle
with GENERATE/CATCH handling.  Workload: mesh 80x160x48 cells (x, y, z) of the
n-FinFET deck, ElectronNumber=2000000, HoleNumber=500000 (plus the reference's
one carrier of each type per silicon cell), dt = 0.1 fs, T = 300 K.

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


def make_workload(refine, electrons, holes, tmp, qc=False):
    inp = emcgpu.write_deck(tmp, refine=refine, electrons=electrons, holes=holes, vd=0.5, vg=0.7,
                            total_step=7000, stat_step=1000, qc=int(qc), quantum_start_step=0)
    return inp


def cpu_baseline(eng, steps):
    """The oracle (CPU restatement, 1 thread) runs complete steps of the same
    workload: deposit, RHS, PCG (warm start from the GPU potential), fields,
    surface rates and the flight of every particle."""
    import oracle_lib
    lib = oracle_lib.load()
    m = eng.mesh
    p = eng.download()
    pot = eng.potential()
    n_ps = 0
    t0 = time.perf_counter()
    for s in range(steps):
        live = int(np.sum(p.i[:p.n] >= 0))
        node_q, _, cell_q, _ = oracle_lib.deposit(eng, p, p.n)
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
        p, _, _ = oracle_lib.fly_step(eng, p, p.n, f, 10 ** 9 + s)
        n_ps += live
    dt = time.perf_counter() - t0
    return n_ps / dt, dt, n_ps


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


def measure_traffic(bench_args):
    """HBM bytes of one flight (k_fly_fast + k_fly_slow) from FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reports for a 512 MiB
    device-to-device copy (the calibration)."""
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

    f_fly, w_fly = per_step(sf), per_step(sw)
    if f_fly is None or w_fly is None:
        return None
    out = {"bytes": (f_fly * fscale + w_fly * wscale) * 1024.0, "fetch_kib": f_fly, "write_kib": w_fly,
           "fetch_scale": fscale, "write_scale": wscale}
    # the flight is latency bound: its L2 hit rate says where the table reads are served
    hm = _pmc_pass("TCC_HIT_sum TCC_MISS_sum", bench_args, os.path.join(tmp, "hm"))
    if hm:
        hit, miss = per_step(hm["TCC_HIT_sum"]), per_step(hm["TCC_MISS_sum"])
        if hit is not None and miss is not None and hit + miss > 0:
            out["l2_hit_rate"] = hit / (hit + miss)
    return out


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--refine", type=int, default=4)
    ap.add_argument("--electrons", type=int, default=2000000)
    ap.add_argument("--holes", type=int, default=500000)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes (roofline.traffic = null)")
    ap.add_argument("--qc", action="store_true",
                    help="quantum correction on (Flag_QuantumCorrection=1 from step 0; not the default workload)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.copy_probe > 0:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    traffic = None
    if not a.no_traffic:
        # before this process touches the GPU: the counter passes run as children
        traffic = measure_traffic(["--steps", "1", "--warmup", "1", "--no-cpu", "--no-traffic",
                                   "--refine", str(a.refine), "--electrons", str(a.electrons),
                                   "--holes", str(a.holes)] + (["--qc"] if a.qc else []))
        print(f"[bench] traffic {traffic}", file=sys.stderr, flush=True)
    tmp = tempfile.mkdtemp(prefix="emc_bench_")
    inp = make_workload(a.refine, a.electrons, a.holes, tmp, qc=a.qc)
    print(f"[bench] deck {inp}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    eng = emcgpu.Engine(inp, device=0, seed=1)
    n0 = eng.particles_init()
    print(f"[bench] {n0} particles on {eng.mesh.cnx}x{eng.mesh.cny}x{eng.mesh.cnz} cells, "
          f"{eng.mesh.nnode} nodes, setup {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    for w in range(a.warmup):
        eng.step(1)
        print(f"[bench] warmup step {w} done", file=sys.stderr, flush=True)
    n_start = eng.count()
    lib = eng.lib
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
    fly_ms = phase[4]
    achieved = FLY_BYTES_PER_PARTICLE * n_avg / (fly_ms * 1e-3) / 1e9 if fly_ms > 0 else None
    print(f"[bench] {a.steps} steps in {elapsed:.3f}s; per step ms: deposit+rhs {phase[0]:.3f} "
          f"pcg {phase[1]:.3f} field {phase[2]:.3f} fly(all) {phase[3]:.3f} k_fly {phase[4]:.3f}; "
          f"pcg iters {np.mean(pcg_it):.1f}", file=sys.stderr, flush=True)
    # MultipleRefresh of the same population, once, after the timed steps (the bench
    # deck, like the reference's plain configuration, runs without it)
    t_mr = time.perf_counter()
    mr_before, mr_after = eng.multiple_refresh(10 ** 6)
    mr_ms = (time.perf_counter() - t_mr) * 1e3
    print(f"[bench] multiple refresh {mr_before} -> {mr_after} particles in {mr_ms:.2f} ms", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu and a.cpu_steps > 0:
        v, dt, nps = cpu_baseline(eng, a.cpu_steps)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"{a.cpu_steps} complete steps of the same workload ({nps} particle-steps, {dt:.1f} s) "
                         f"by the oracle in C on one host core"}
        print(f"[bench] cpu baseline {v:.4g} particle-steps/s over {dt:.1f}s", file=sys.stderr, flush=True)
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
                               f"{n_start} particles (ElectronNumber={a.electrons}, HoleNumber={a.holes}), "
                               f"dt 0.1 fs, full step: deposit + PCG Poisson + fields + flight"
                               + (" + quantum-correction potential" if a.qc else ""),
                   "quantum_correction": bool(a.qc),
                   "particles": int(n_start), "nodes": int(eng.mesh.nnode),
                   "phase_ms": {"deposit_rhs": phase[0], "poisson": phase[1], "field": phase[2],
                                "flight_total": phase[3], "k_fly": phase[4]},
                   "pcg_iters_mean": float(np.mean(pcg_it)),
                   "multiple_refresh_ms": mr_ms,
                   "full_event_loop_fraction": full_loop / max(1, n_ps),
                   "events_per_particle_step": events / max(1, n_ps)},
        "roofline": {"bound": "hbm", "kernel": "k_fly_fast + k_fly_slow", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_counters": traffic,
                     "l2_hit_rate": traffic.get("l2_hit_rate") if traffic else None,
                     "bytes_per_particle": FLY_BYTES_PER_PARTICLE,
                     "algorithmic_bytes": FLY_BYTES_PER_PARTICLE * n_avg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
