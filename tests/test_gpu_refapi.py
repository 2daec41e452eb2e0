"""A case file written against the reference's headers (lbm3_amd/refapi/
cases/sphereFacade/sphereFacadeMain.cu: hooks compiled with the case,
buildGrids, updateAllGrids, the tracker, fArray views and a TNL reduce) runs
on the GPU and equals the same case driven through the C ABI with rule
tables (lib.Sim): macroscopic fields and the per-step wall-force tracker bit
for bit, kinetic energy to the summation order, section cut bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from lbm3_amd import cases, geometry, lib

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "lbm3_amd", "refapi", "cases", "sphereFacade", "sphereFacade")
STEPS = 30


def _read_log(path):
    out = {"force": []}
    for line in open(path):
        k, *v = line.split()
        if k == "force":
            out["force"].append([float(x) for x in v])
        else:
            out[k] = float(v[0])
    return out


def test_case_file_equals_the_c_abi_path(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("the case executable is not built (__graft_entry__.build)")
    case, dom, tris = cases.sphere(48, 24, 24, diameter=6.0, center=(14.0, 12.0, 12.0), segments=(16, 32))
    stl = tmp_path / "sphere.stl"
    geometry.write_binary_stl(stl, tris[0])
    prefix = str(tmp_path / "out")
    print("running %s %d steps" % (EXE, STEPS), flush=True)
    r = subprocess.run([EXE, str(stl), str(STEPS), prefix], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = _read_log(prefix + ".txt")

    case.trackOpenBoundaries = 1
    s = lib.Sim(case, "f32")
    fu = s.build(dom, tris)
    s.init()
    forces = []
    for _ in range(STEPS):
        s.run(1)
        w = s.update_tracker(walls=1)["walls"][0]
        forces.append(np.float32(w[:3]))
    s.sync()
    assert log["fluid_updates"] == fu and log["iterations"] == STEPS
    got = np.fromfile(prefix + ".macros", np.float32).reshape(-1, 4)
    want = s.macros(0)
    print("macros: %d cells, max |case - C ABI| %.3e" % (len(want), np.abs(got - want).max()))
    assert got.shape == want.shape and np.array_equal(got, want)
    f_case = np.array([x[1:4] for x in log["force"]], np.float32)
    assert np.array_equal(f_case, np.array(forces)), (f_case[:3], forces[:3])
    ke = s.kinetic_energy()
    print("kinetic energy: case %.15e, C ABI %.15e" % (log["kinetic_energy"], ke))
    assert abs(log["kinetic_energy"] / ke - 1.0) < 1e-12
    raw = np.fromfile(prefix + ".cut", np.int32)
    V, H = raw[1], raw[2]
    img = raw[4:].view(np.float32).reshape(V, H, 6)
    assert raw[0] == 1 and raw[3] == 6 and np.array_equal(img, s.section_cut("xy", 12.0))
    cd = -2.0 * forces[-1][0] / (1.0 * 1.0 * 3.14159 * 0.006 ** 2 / 4.0) / 1.0
    print("drag coefficient after %d steps: %.4f (case file %.4f)" % (STEPS, cd, log["force"][-1][4]))
    s.close()
