"""Writes finfet_fields6999.npz: transport observables of the reference's own run
of its shipped FinFET deck, taken from the result files it ships (data only: the
expected outputs of the whole-run statistical test).

- heat: finfet/heat6999.vtr, the node field -(electron + hole heat column) of
  data/heat6999 as finfet/scripts/plot_3D_heat.py writes it (the heating-phase
  window of steps 6000..6999), float32 [i][j][k];
- electron: finfet/Electron6999_realspace.vtr, the per-node sum over the rows of
  data/Electron6999 of column 7 (n), as finfet/scripts/plot_3D_realspace.py
  writes it (the file holds the transient window and the heating-phase window
  of the same number, appended), float32 [i][j][k];
- heat_profile_y: finfet/heat_profile_y.csv (y [um], W/m^3), the y profile of
  finfet/scripts/plot_1D_heat_profile.py;
- contact currents per 1000-step window (A), read off the reference's plots of
  data/current (finfet/scripts/plot_current.py draws one marker per record):
  * ids_vs_step_contact1 [7][2]: finfet/ids_vs_step.png, contact 1, the file of
    a run from the sampled population: its 7 transient windows and its 7
    heating-phase windows drawn over the same step axis (the polyline runs from
    the transient's step 7000 back to step 1000).  At steps 1000 and 2000 the two
    markers are apart (transient 7.79e-12 and 3.73e-12 above, heating phase
    below); from step 3000 on they overlap and are given as [upper, lower],
    unassigned.
  * current_contact1 [7]: finfet/Idrain.png, contact 1, one series of 7 windows.
  * current_contact0 [7]: finfet/ids_vs_step2.png, contact 0, one series of 7
    windows (the same values as finfet/Isource.png, on a finer axis).
  Each marker is located from the pixels: its centre along the column of its
  step's vertical grid line (two overlapping markers: the ends of the blue run
  less half a marker diameter), converted to current by a straight-line fit
  through the horizontal grid lines and their tick values.  One pixel is
  5.3e-15 A on ids_vs_step.png, 1.1e-15 A on Idrain.png and 4.6e-17 A on
  ids_vs_step2.png.

    python tests/golden/make_shipped_fields.py /root/reference
"""
import os
import re
import sys

import numpy as np


def read_vtr(path):
    text = open(path).read()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', text).group(1).split()]
    nx, ny, nz = ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1
    block = re.search(r"<DataArray[^>]*>(.*?)</DataArray>", text, re.S).group(1)
    return np.array(block.split(), dtype=np.float32).reshape(nz, ny, nx).transpose(2, 1, 0).copy()


def _groups(idx):
    """Runs of consecutive indices -> (first, last) pairs."""
    out = []
    for v in idx:
        if out and v == out[-1][1] + 1:
            out[-1][1] = v
        else:
            out.append([v, v])
    return out


def read_markers(path, ticks):
    """Marker values of a one-colour matplotlib line plot with a grid.
    `ticks` are the values of the interior horizontal grid lines, top to bottom.
    Returns one list per vertical grid line (left to right) of the values of the
    markers centred on it, highest first."""
    from PIL import Image
    im = np.asarray(Image.open(path).convert("RGB")).astype(int)
    H, W, _ = im.shape
    r, g, b = im[:, :, 0], im[:, :, 1], im[:, :, 2]
    gray = (abs(r - g) < 4) & (abs(g - b) < 4) & (r > 200) & (r < 245)
    # interior grid lines are 3 pixels wide; the frame edges are single rows
    rows = [(a + z) / 2 for a, z in _groups(np.nonzero(gray.sum(axis=1) > W * 0.5)[0]) if z - a == 2]
    cols = [(a + z) / 2 for a, z in _groups(np.nonzero(gray.sum(axis=0) > H * 0.5)[0]) if z - a == 2]
    assert len(rows) == len(ticks), (path, rows)
    slope, icpt = np.polyfit(rows, ticks, 1)
    blue = (abs(r - 31) < 40) & (abs(g - 119) < 40) & (abs(b - 180) < 40)  # matplotlib C0
    runs = []
    for x in cols:
        col = blue[:, int(round(x))]
        runs.append([(a, z) for a, z in _groups(np.nonzero(col)[0]) if z - a > 20])
    diam = min(z - a for rr in runs for a, z in rr) + 1
    out = []
    for rr in runs:
        centres = []
        for a, z in rr:
            if z - a + 1 <= diam + 1:
                centres.append((a + z) / 2)
            else:  # two overlapping markers
                centres += [a + (diam - 1) / 2, z - (diam - 1) / 2]
        out.append(sorted((slope * y + icpt for y in centres), reverse=True))
    return out


def main(ref):
    fin = os.path.join(ref, "finfet")
    heat = read_vtr(os.path.join(fin, "heat6999.vtr"))
    elec = read_vtr(os.path.join(fin, "Electron6999_realspace.vtr"))
    prof = np.loadtxt(os.path.join(fin, "heat_profile_y.csv"), delimiter=",", skiprows=1)
    both = read_markers(os.path.join(fin, "ids_vs_step.png"), [7e-12, 6e-12, 5e-12, 4e-12])
    assert all(len(m) == 2 for m in both), both
    c1 = read_markers(os.path.join(fin, "Idrain.png"), [3.8e-12, 3.6e-12, 3.4e-12, 3.2e-12])
    c0 = read_markers(os.path.join(fin, "ids_vs_step2.png"), [-3.430e-12 - 0.005e-12 * q for q in range(9)])
    assert all(len(m) == 1 for m in c1 + c0)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "finfet_fields6999.npz")
    np.savez_compressed(out, heat=heat, electron=elec, heat_profile_y=prof.astype(np.float64),
                        ids_vs_step_contact1=np.array(both), current_contact1=np.array([m[0] for m in c1]),
                        current_contact0=np.array([m[0] for m in c0]))
    print("ids_vs_step contact 1:", np.array(both) * 1e12)
    print("Idrain contact 1:", np.array(c1).ravel() * 1e12)
    print("ids_vs_step2 contact 0:", np.array(c0).ravel() * 1e12)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
