"""Build the Fortran test driver of the nekamd shims with flang.

The interface module and the shims (nek5000_amd/fortran/) are compiled with
Nek5000's flags (fixed form, REAL and DOUBLE PRECISION of 8 bytes,
makenek.inc:237-270) against the stand-in SIZE/TOTAL of tests/fortran/stub,
with the stand-in Nek routines of nek_stubs.f, and linked with
libnekamd.so.  Output: tests/fortran/_build/drive.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "nek5000_amd", "fortran")
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "drive")
LIBDIR = os.path.join(ROOT, "nek5000_amd")

FFLAGS = ["-O2", "-ffixed-form", "-fdefault-real-8", "-fdefault-double-8", "-cpp"]


def flang():
    for c in ("flang", "amdflang", "/opt/rocm/llvm/bin/flang", "/opt/rocm/lib/llvm/bin/flang"):
        p = shutil.which(c) or (c if os.path.exists(c) else None)
        if p:
            return p
    return None


def build(verbose=False):
    fc = flang()
    if fc is None:
        raise RuntimeError("no flang compiler")
    os.makedirs(OUT, exist_ok=True)
    inc = ["-I", os.path.join(HERE, "stub"), "-module-dir", OUT, "-I", OUT]
    objs = []
    for src in (os.path.join(SRC, "nekamd_mod.f"), os.path.join(SRC, "nekamd_shim.f"),
                os.path.join(HERE, "nek_stubs.f"), os.path.join(HERE, "drive.f")):
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        cmd = [fc] + FFLAGS + inc + ["-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed for {src}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    # the library is found next to the executable's tree wherever it is copied
    rpath = "-Wl,-rpath,$ORIGIN/../../../nek5000_amd"
    cmd = [fc, "-o", EXE] + objs + ["-L", LIBDIR, "-lnekamd", rpath]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", EXE)
    return EXE


if __name__ == "__main__":
    build(verbose=True)
