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
- drain current per 1000-step transient window, steps 999 .. 6999: contact 1
  and contact 0 in A, read off finfet/Idrain.png (contact 1, axis 3.0-4.0e-12,
  read to +-0.005e-12) and finfet/multi_current.png (contact 0, read to
  +-0.01e-12).

    python tests/golden/make_shipped_fields.py /root/reference
"""
import os
import re
import sys

import numpy as np

# read off the reference's plots (Idrain.png, multi_current.png), steps 1000 .. 7000
CURRENT_CONTACT1 = [3.420e-12, 3.409e-12, 3.420e-12, 3.427e-12, 3.441e-12, 3.460e-12, 3.440e-12]
CURRENT_CONTACT0 = [-3.45e-12, -3.44e-12, -3.455e-12, -3.46e-12, -3.465e-12, -3.465e-12, -3.47e-12]


def read_vtr(path):
    text = open(path).read()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', text).group(1).split()]
    nx, ny, nz = ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1
    block = re.search(r"<DataArray[^>]*>(.*?)</DataArray>", text, re.S).group(1)
    return np.array(block.split(), dtype=np.float32).reshape(nz, ny, nx).transpose(2, 1, 0).copy()


def main(ref):
    fin = os.path.join(ref, "finfet")
    heat = read_vtr(os.path.join(fin, "heat6999.vtr"))
    elec = read_vtr(os.path.join(fin, "Electron6999_realspace.vtr"))
    prof = np.loadtxt(os.path.join(fin, "heat_profile_y.csv"), delimiter=",", skiprows=1)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "finfet_fields6999.npz")
    np.savez_compressed(out, heat=heat, electron=elec, heat_profile_y=prof.astype(np.float64),
                        current_contact1=np.array(CURRENT_CONTACT1), current_contact0=np.array(CURRENT_CONTACT0))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
