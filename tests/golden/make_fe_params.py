"""Writes fe_params.txt: the electron scattering-factor parameters of
Kirkland (Advanced Computing in Electron Microscopy, 2nd ed., Springer 2010,
Appendix C), 12 per element for Z = 1..103 (three Lorentzians a_i, b_i then
three Gaussians c_i, d_i), as the reference hard-codes them in
TemsimLibrary/slicelib.cpp:958-2209 (ReadfeTable).

The published table is data the multislice model needs as is; this script
reads it from the reference's source text, in this container, and writes
it one element per line.  Run once; the output is committed."""
import os
import re
import sys

SRC = "/root/reference/TemsimLibrary/slicelib.cpp"


OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                   "utr_public_amd", "data", "fe_params.txt")


def main(src=SRC, out=OUT):
    text = open(src).read()
    vals = {}
    for z, i, v in re.findall(r"fparams\[(\d+)\]\[(\d+)\]\s*=\s*([-+0-9.eE]+)\s*;", text):
        vals[(int(z), int(i))] = v
    rows = []
    for z in range(1, 104):
        rows.append(" ".join("%.8e" % float(vals[(z, i)]) for i in range(12)))
    with open(out, "w") as f:
        f.write("# Kirkland (2010) Appendix C electron scattering-factor parameters,\n")
        f.write("# one line per Z = 1..103: a1 b1 a2 b2 a3 b3 (Lorentzians) c1 d1 c2 d2 c3 d3 (Gaussians)\n")
        f.write("\n".join(rows) + "\n")
    print("wrote", out, len(rows), "elements")


if __name__ == "__main__":
    main(*sys.argv[1:])
