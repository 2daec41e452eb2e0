"""The case-facing header layer (lbm3_amd/refapi/include): reference case
files compile against it unchanged for gfx950 (hipcc on the CPU, no GPU).

The reference's own cases are compiled where the reference tree is present
(they are read from it, copied into a temporary build tree, never edited);
the package's own reference-style case compiles and links everywhere."""
import os

import pytest

from lbm3_amd.refapi import build_case

REF = "/root/reference/cases"
REF_CASES = ["taylorGreenVortex/taylorGreenVortexMain.cu", "flowAroundSphere/flowAroundSphereMain.cu",
             "poiseuilleFlow/poiseuilleFlowMain.cu", "torqueTest/torqueTestMain.cu",
             "nonReflectiveTest/nonReflectiveTestMain.cu", "ns300/ns300Main.cu", "BruteforcePumpOnly/BruteforcePump.cu"]


@pytest.mark.parametrize("case", REF_CASES)
def test_reference_case_compiles_unchanged(tmp_path, case):
    src = os.path.join(REF, case)
    if not os.path.exists(src):
        pytest.skip("reference case files are not present")
    before = open(src, "rb").read()
    r = build_case.build(src, str(tmp_path / "case.o"), compile_only=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.getsize(tmp_path / "case.o") > 0
    assert open(src, "rb").read() == before


def test_package_case_links(tmp_path):
    src = os.path.join(os.path.dirname(build_case.HERE), "refapi", "cases", "sphereFacade", "sphereFacadeMain.cu")
    r = build_case.build(src, str(tmp_path / "sphereFacade"))
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.access(tmp_path / "sphereFacade", os.X_OK)
