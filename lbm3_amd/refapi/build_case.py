"""Compile an LBM3 case file, unchanged, against the MI355X library.

A reference case file includes its headers as "../../include/<name>.h" from
cases/<case>/ (taylorGreenVortexMain.cu:46-49).  This script copies the file
into a build tree shaped like the reference repository, whose include/ is the
case-facing header layer of this package (lbm3_amd/refapi/include), and
compiles it with hipcc for gfx950, linked to liblbm3mi.so.

    python -m lbm3_amd.refapi.build_case path/to/caseMain.cu [-o executable]

The case file itself is not edited: a case that drops in unchanged compiles
here as it is.  Case files written for this package can also live directly in
lbm3_amd/refapi/cases/<case>/, next to the include/ they name.
"""
import argparse
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-x", "hip", "-O3", "-std=c++17", "--offload-arch=gfx950", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def command(source, output, compile_only=False):
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc")]
    if compile_only:
        return cmd + ["-c", source, "-o", output]
    return cmd + [source, "-o", output, "-L", PKG, "-llbm3mi", "-Wl,-rpath," + PKG]


def build(case_file, output, compile_only=False, verbose=False):
    """Stage `case_file` as <tmp>/cases/<case>/<file> beside <tmp>/include (the
    header layer) and compile it; returns the compiler's CompletedProcess."""
    case_file = os.path.abspath(case_file)
    name = os.path.basename(os.path.dirname(case_file)) or "case"
    with tempfile.TemporaryDirectory(prefix="lbm3case") as tmp:
        os.makedirs(os.path.join(tmp, "cases", name))
        staged = os.path.join(tmp, "cases", name, os.path.basename(case_file))
        shutil.copyfile(case_file, staged)
        os.symlink(os.path.join(HERE, "include"), os.path.join(tmp, "include"))
        cmd = command(staged, os.path.abspath(output), compile_only)
        if verbose:
            print(" ".join(cmd))
        return subprocess.run(cmd, capture_output=True, text=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("case_file")
    ap.add_argument("-o", "--output", default=None)
    ap.add_argument("-c", "--compile-only", action="store_true", help="compile to an object file, do not link")
    a = ap.parse_args(argv)
    out = a.output or os.path.splitext(os.path.basename(a.case_file))[0] + (".o" if a.compile_only else "")
    r = build(a.case_file, out, a.compile_only, verbose=True)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    return r.returncode


if __name__ == "__main__":
    sys.exit(main())
