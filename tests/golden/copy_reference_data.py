"""Copy the data files of the reference's own runs that the tests compare with
into this folder as fixtures.  Only data is copied (numbers and one STL mesh);
nothing of the reference is run.

* Taylor-Green kinetic-energy curves (cases/taylorGreenVortex/results/
  *_kineticEnergy.txt, one "time star;normalized kinetic energy" row per
  sample, written by getKineticEnergy, taylorGreenVortexMain.cu:116-180):
  COARSE 128^3 uniform, FINE 256^3 uniform, QUADRATIC 128^3 with the central
  cube refined once, LINEAR the same with the linear interface
  (INTERFACE_VERSIONS_BACKUP/updateInterfaceLINEAR.h); Re = 1600, Ma = 0.1,
  0 <= t* <= 20.
* Inlet histories of the non-reflective pipe test (cases/nonReflectiveTest/
  {characteristic,dirichlet,impedance2e-5,impedance2e-6,impedance2e-7}.txt,
  one row per iteration: iteration, area-averaged inlet velocity [m/s],
  area-averaged inlet pressure [Pa], relative L2 error of the inlet velocity
  profile; nonReflectiveTestMain.cu:208-233): the characteristic
  non-reflective inlet, the strict Dirichlet inlet and the impedance scheme
  (updateGrid_SCHLAFFER_IMPEDANCE_SCHEME.h) at three density tolerances.
  Rows 1-2000 are kept whole, then every 100th row up to 50000.
* Centre-line pressure of the Poiseuille pipe case after 20000 iterations
  (cases/poiseuilleFlow/results/INTERFACE/{COARSE,FINE}_centerlinePressure.txt,
  "z [mm];pressure [Pa]" along the pipe axis from the x = 0 section cut
  poiseuilleFlowMain.cu:158-168, exportCenterlinePressure.py): uniform grids
  of 0.4 mm and 0.2 mm cells, and 0.4 mm cells with the core refined once
  (the quadratic interface).
* The pipe mesh both pipe cases read (pipeD20L200STL.STL, binary STL,
  376 triangles).

    python tests/golden/copy_reference_data.py [reference root]
"""
import os
import shutil
import sys

TGV = ("COARSE", "FINE", "QUADRATIC", "LINEAR")
INLET = ("characteristic", "dirichlet", "impedance2e-5", "impedance2e-6", "impedance2e-7")
POISEUILLE = ("COARSE", "FINE", "QUADRATIC")


def _rows(path):
    with open(path) as f:
        rows = f.read().splitlines()
    return rows[0], [r for r in rows[1:] if r.strip()]


def main(root):
    here = os.path.dirname(os.path.abspath(__file__))
    for name in TGV:
        header, body = _rows(os.path.join(root, "cases", "taylorGreenVortex", "results", "%s_kineticEnergy.txt" % name))
        for r in body:                                 # every row is two numbers
            t, e = r.split(";")
            float(t), float(e)
        with open(os.path.join(here, "tgv_%s_kinetic_energy.txt" % name.lower()), "w") as f:
            f.write(header + "\n" + "\n".join(body) + "\n")
        print(name, len(body), "rows")
    for name in INLET:
        header, body = _rows(os.path.join(root, "cases", "nonReflectiveTest", "%s.txt" % name))
        keep = []
        for r in body:
            it = int(r.split(";")[0])
            if it <= 2000 or it % 100 == 0:
                keep.append(r)
        with open(os.path.join(here, "pipe_inlet_%s.txt" % name), "w") as f:
            f.write(header + "\n" + "\n".join(keep) + "\n")
        print(name, len(keep), "of", len(body), "rows")
    for name in POISEUILLE:
        header, body = _rows(os.path.join(root, "cases", "poiseuilleFlow", "results", "INTERFACE",
                                          "%s_centerlinePressure.txt" % name))
        with open(os.path.join(here, "poiseuille_%s_centerline_pressure.txt" % name.lower()), "w") as f:
            f.write(header + "\n" + "\n".join(body) + "\n")
        print(name, len(body), "rows")
    shutil.copyfile(os.path.join(root, "cases", "nonReflectiveTest", "pipeD20L200STL.STL"),
                    os.path.join(here, "pipeD20L200STL.STL"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
