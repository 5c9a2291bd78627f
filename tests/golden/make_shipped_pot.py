"""Writes finfet_pot6999.npz: the window-mean potential [V] at step 6999 of the
reference's own run of its shipped FinFET deck (finfet/input.txt, ldg.txt,
lgrid.txt), taken from the result file it ships, finfet/pot6999.vtr (VTK
rectilinear grid, x fastest, then y, then z).  Stored as float32 [k][j][i].
Data only: the expected output of the whole-run statistical test.

    python tests/golden/make_shipped_pot.py /root/reference
"""
import os
import re
import sys

import numpy as np


def main(ref):
    text = open(os.path.join(ref, "finfet", "pot6999.vtr")).read()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', text).group(1).split()]
    nx, ny, nz = ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1
    block = re.search(r'Name="potential"[^>]*>(.*?)</DataArray>', text, re.S).group(1)
    pot = np.array(block.split(), dtype=np.float32).reshape(nz, ny, nx)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "finfet_pot6999.npz")
    np.savez_compressed(out, potential=pot)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
