"""The FAST-side binding (fast_amd/binding/ctrlgpu.cpp) compiles against the
reference's own headers and include/fgpu.h, with the reference's build flags
(the same g++ line as oracle/Makefile), and its symbols resolve against the
reference's objects and libfgpu.so.  Skipped where the reference sources are
absent."""
import os
import subprocess

import pytest

from conftest import ROOT

REF = "/root/reference"
SUBDIRS = ["ctrl", "elc", "geom", "load", "main", "mat", "mlib", "node", "sec", "shp",
           "spring", "heat"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not present")
def test_binding_compiles_against_reference_headers(tmp_path):
    inc = ["-I" + os.path.join(REF, d) for d in SUBDIRS]
    obj = str(tmp_path / "ctrlgpu.o")
    cmd = ["g++", "-O1", "-w", "-D_UNIX_", "-D_DEBUG_"] + inc + [
        "-I" + os.path.join(ROOT, "include"), "-c",
        os.path.join(ROOT, "fast_amd", "binding", "ctrlgpu.cpp"), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
    # the binding uses the reference's interfaces and the C-ABI, nothing else
    nm = subprocess.run(["nm", "-C", "--undefined-only", obj], capture_output=True,
                        text=True).stdout
    assert "fgpu_create" in nm and "fgpu_assemble_K" in nm and "fgpu_int_force" in nm
    assert "cPrescDof::GetPrescDofs" in nm and "fgpu_set_penalty" in nm
    assert "cBspPat::getExtMat" in nm


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ref_gpu")),
                    reason="the reference build (oracle/_ref) is not present")
def test_binding_driver_is_linked():
    """oracle/_ref/ref_gpu (the reference's set-up + the binding) links: every
    symbol the binding takes from the reference and from libfgpu.so resolves."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_gpu")
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libfgpu.so" in out and "not found" not in out
