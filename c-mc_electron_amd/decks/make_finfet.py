#!/usr/bin/env python3
"""Write input.txt / ldg.txt / lgrid.txt decks for the n-FinFET workloads.

The decks use the reference's deck grammar (finfet/ldg_spec.md, parsed by
src/mcmodel.cpp:4190-4584) and axis convention (x = fin height, y = transport,
z = fin width; finfet/ldg.txt:6-27).

Device (all positions in nm, grid lines every 1/refine nm):
  x: gate oxide [-1,0] | Si fin [0,10] | Si substrate [10,19]
  y: source [-20,-7] | gate [-7,8] (Lg = 15 nm) | drain [8,20]
  z: vacuum [-3,-1] | oxide [-1,0] | fin [0,5] | oxide [5,6] | vacuum [6,9]
The 40 x 1 nm y-grid cannot hold 10|15|10 nm exactly, so the source/drain are
13 nm / 12 nm long including their one-cell contact reservoirs.

Usage: make_finfet.py OUTDIR --refine R --electrons NE --holes NH [--vd V --vg V] [--lg L --wfin W]
  refine=1 gives 20x40x12 cells (x, y, z); refine=4 gives 80x160x48 cells.
  --lg / --wfin set the gate length and fin width in whole nm (default 15 and 5);
  y then runs over [-20, lg - 7 + 12] and z over [-3, wfin + 4].
"""
import argparse
import os


def grid_axis(lo, hi, refine):
    n = (hi - lo) * refine
    return [(lo + i / refine) * 1e-3 for i in range(n + 1)]  # um


def write_decks(out, refine=1, electrons=20000, holes=2000, vs=0.0, vd=0.5, vg=0.7,
                total_step=2000, dt=1e-16, stat_step=1000, heat=0, heat_steps=0,
                surface=1, qc=0, quantum_start_step=20, lg=15, wfin=5, phims=-0.46, mr=0, mr_step=10):
    """lg: gate length and wfin: fin width, whole nm (defaults 15 and 5: the
    device above); the source stays 13 nm and the drain 12 nm long.  phims: the
    gate's work-function difference, the second number of the gate contact (the
    reference's FinFET decks use -0.46 V).  mr, mr_step: Flag_MultipleRefresh
    and its period."""
    os.makedirs(out, exist_ok=True)
    y0, g0 = -20, -7
    g1 = g0 + lg
    y1 = g1 + 12
    w, w1, zt = wfin, wfin + 1, wfin + 4
    axes = [grid_axis(-1, 19, refine), grid_axis(y0, y1, refine), grid_axis(-3, zt, refine)]
    with open(os.path.join(out, "lgrid.txt"), "w") as f:
        for ax in axes:
            f.write("%d\n" % len(ax))
            for v in ax:
                f.write("%.10g\n" % v)
    r = 1.0 / refine  # one cell in nm
    ldg = f"""default_par_number 1 1
region -1 19 {y0} {y1} -3 {zt} VACUUM
region 0 10 {y0} {y1} 0 {w} SILICON
region 10 19 {y0} {y1} -3 {zt} SILICON
region -1 0 {y0} {y1} -1 {w1} OXIDE
region 0 10 {y0} {y1} -1 0 OXIDE
region 0 10 {y0} {y1} {w} {w1} OXIDE
donor -1 19 {y0} {y1} -3 {zt} 0
acceptor -1 19 {y0} {y1} -3 {zt} 0
donor 0 10 {y0} {g0} 0 {w} 1e26
donor 0 10 {g1} {y1} 0 {w} 1e26
acceptor 0 10 {g0} {g1} 0 {w} 1e21
acceptor 10 19 {y0} {y1} -3 {zt} 1e21
motioncube -1 19 {y0} {y1} -3 {zt} PASS PASS PASS PASS PASS PASS
motionplane 0 10 {y0 + r:g} {y0 + r:g} 0 {w} LEFT CATCH
motionplane 0 10 {y0 + r:g} {y0 + r:g} 0 {w} RIGHT GENERATE
motionplane 0 10 {y0} {y0} 0 {w} LEFT REFLECT
motionplane 0 10 {y1 - r:g} {y1 - r:g} 0 {w} RIGHT CATCH
motionplane 0 10 {y1 - r:g} {y1 - r:g} 0 {w} LEFT GENERATE
motionplane 0 10 {y1} {y1} 0 {w} RIGHT REFLECT
motionplane 0 0 {y0} {y1} 0 {w} UP SCATTOX
motionplane 0 10 {y0} {y1} 0 0 FRONT SCATTOX
motionplane 0 10 {y0} {y1} {w} {w} BACK SCATTOX
motionplane 10 10 {y0} {y1} -3 0 UP SCATTOX
motionplane 10 10 {y0} {y1} {w} {zt} UP SCATTOX
motionplane 10 19 {y0} {y0} -3 {zt} LEFT SCATTOX
motionplane 10 19 {y1} {y1} -3 {zt} RIGHT SCATTOX
motionplane 10 19 {y0} {y1} -3 -3 FRONT SCATTOX
motionplane 10 19 {y0} {y1} {zt} {zt} BACK SCATTOX
motionplane {19 - r:g} {19 - r:g} {y0} {y1} -3 {zt} UP GENERATE
motionplane {19 - r:g} {19 - r:g} {y0} {y1} -3 {zt} DOWN CATCH
motionplane 19 19 {y0} {y1} -3 {zt} DOWN SCATTOX
attachcontact 0 10 {y0} {y1} 0 {w} 0
attachcontact 0 10 {y0} {y0 + 2 * r:g} 0 {w} 1
attachcontact 0 10 {y1 - 2 * r:g} {y1} 0 {w} 2
attachcontact {19 - 2 * r:g} 19 {y0} {y1} -3 {zt} 3
ScatterArea 0 10 {y0} {y1} 0 {w} 3
ep_parm 4 0.5 4 0.5 4 0.5 3.1
surfaces 3
  0 0 0
  2 0 4
  2 {w} 5
surface_scatter_range 0 10 {y0} {y1} 0 {w}
contact
  1 0
  0 10 {y0} {y0} 0 {w}
  {vs:g}
contact
  1 0
  0 10 {y1} {y1} 0 {w}
  {vd:g}
contact
  1 0
  19 19 {y0} {y1} -3 {zt}
  0
contact
  3 {phims:g}
  -1 -1 {g0} {g1} -1 {w1}
  -1 10 {g0} {g1} -1 -1
  -1 10 {g0} {g1} {w1} {w1}
  {vg:g}
end
"""
    with open(os.path.join(out, "ldg.txt"), "w") as f:
        f.write(ldg)
    inp = f"""gridFile=lgrid.txt
device_file_name=ldg.txt
total_step={total_step}
ElectronNumber={electrons}
HoleNumber={holes}
Temperature=300
dt={dt:g}
stat_step={stat_step}
stat_heat_step={stat_step}
heat_steps={heat_steps}
Flag_QuantumCorrection={qc}
quantum_start_step={quantum_start_step}
Flag_MultipleRefresh={mr}
Flag_compute_heat={heat}
Flag_SurfaceScatter={surface}
mr_step={mr_step}
"""
    with open(os.path.join(out, "input.txt"), "w") as f:
        f.write(inp)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--electrons", type=int, default=20000)
    ap.add_argument("--holes", type=int, default=2000)
    ap.add_argument("--vd", type=float, default=0.5)
    ap.add_argument("--vg", type=float, default=0.7)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--heat", type=int, default=0)
    ap.add_argument("--heat-steps", type=int, default=0)
    ap.add_argument("--qc", type=int, default=0)
    ap.add_argument("--lg", type=int, default=15)
    ap.add_argument("--wfin", type=int, default=5)
    a = ap.parse_args()
    write_decks(a.out, a.refine, a.electrons, a.holes, vd=a.vd, vg=a.vg, total_step=a.steps,
                heat=a.heat, heat_steps=a.heat_steps, qc=a.qc, lg=a.lg, wfin=a.wfin)


if __name__ == "__main__":
    main()
