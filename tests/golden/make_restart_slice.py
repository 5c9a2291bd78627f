"""Writes par_info_restart_slice.txt: the rows of cells i < 20, j < 10, k < 12
of the reference's restart file, then its "-1 -1 -1" terminator,
finfet_old/input/finfet-uniform/par_info_restart_vg1_vd1_noqc_99999 (a file the
reference's output_for_restart wrote, src/mcmodel.cpp:895-965), unchanged and in
file order.  Data only: the rows are inputs and expected output of the
restart writer's format test.

    python tests/golden/make_restart_slice.py /root/reference
"""
import os
import sys

SRC = "finfet_old/input/finfet-uniform/par_info_restart_vg1_vd1_noqc_99999"


def main(ref):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "par_info_restart_slice.txt")
    with open(os.path.join(ref, SRC)) as f, open(out, "w") as g:
        for line in f:
            v = line.split()
            if (len(v) == 7 and int(v[0]) < 20 and int(v[1]) < 10 and int(v[2]) < 12) or int(v[0]) < 0:
                g.write(line)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
