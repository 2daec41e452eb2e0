"""Copy the reference's Taylor-Green kinetic-energy curves into this folder as
data fixtures (cases/taylorGreenVortex/results/*_kineticEnergy.txt: one
"time star;normalized kinetic energy" row per sample, written by the case's
getKineticEnergy, taylorGreenVortexMain.cu:116-180, every TRACKER_PERIOD
iterations).  Only the rows are copied; nothing of the reference is run.

COARSE: 128^3 uniform grid, FINE: 256^3 uniform, QUADRATIC: 128^3 with the
central cube refined once (the interface as it is in the reference's sources);
Re = 1600, Ma = 0.1, 0 <= t* <= 20 (taylorGreenVortexMain.cu:8-17).

    python tests/golden/copy_tgv_energy.py [reference root]
"""
import os
import sys

NAMES = ("COARSE", "FINE", "QUADRATIC")


def main(root):
    src = os.path.join(root, "cases", "taylorGreenVortex", "results")
    here = os.path.dirname(os.path.abspath(__file__))
    for name in NAMES:
        with open(os.path.join(src, "%s_kineticEnergy.txt" % name)) as f:
            rows = f.read().splitlines()
        header, body = rows[0], [r for r in rows[1:] if r.strip()]
        for r in body:                               # every row is two numbers
            t, e = r.split(";")
            float(t), float(e)
        with open(os.path.join(here, "tgv_%s_kinetic_energy.txt" % name.lower()), "w") as f:
            f.write(header + "\n" + "\n".join(body) + "\n")
        print(name, len(body), "rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
