"""The drop-ins that use the reference's module state (ofd_refmod.f90:
assembleQuenching, interpolation_tet, genSparStruc, assembleNavSto,
setBCs_overset) compile against the reference's own
modules (aAdjKeep, commonvars, commonpars, as built under oracle/_ref) and
export the link names the reference's callers resolve to.  Compile-only:
running them needs the reference's program and a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "oversetdevfort_amd", "fortran")
REFMOD = os.path.join(ROOT, "oracle", "_ref")


def test_refmod_compiles_against_reference_modules(tmp_path):
    if shutil.which("amdflang") is None:
        pytest.skip("amdflang not available")
    if not os.path.exists(os.path.join(REFMOD, "aadjkeep.mod")):
        pytest.skip("reference modules not built (make -C oracle ref)")
    for src, extra in (("ofd_iface.f90", []), ("ofd_refapi.f90", []),
                       ("ofd_refmod.f90", ["-I", REFMOD])):
        subprocess.run(["amdflang", "-O2", "-c", os.path.join(FDIR, src), "-J",
                        str(tmp_path), "-o", str(tmp_path / src.replace(".f90", ".o"))]
                       + extra, check=True, cwd=tmp_path)
    syms = subprocess.run(["nm", str(tmp_path / "ofd_refmod.o")], check=True,
                          capture_output=True, text=True).stdout.split()
    ref = subprocess.run(["nm"] + [os.path.join(REFMOD, f) for f in
                                   ("solveQuenching.o", "interpolation.o", "genSparStruc.o",
                                    "assembleNavSto.o", "setBCs.o")], check=True,
                         capture_output=True, text=True).stdout.split()
    for name in ("assemblequenching_", "interpolation_tet_", "gensparstruc_",
                 "assemblenavsto_", "setbcs_overset_"):
        assert name in syms, name
        assert name in ref, name
