"""Build the Fortran host program host/_build/sac_vac with flang.

The program links sac_amd/libsac_amd.so (rpath to that directory), so the
HIP library must be built first.  Objects and module files go to
host/_build/ (git-ignored); the executable travels with the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "sac_vac")
LIBDIR = os.path.join(ROOT, "sac_amd")
SOURCES = ["sac_bind.f90", "sac_vac.f90"]


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang"),
                 shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("flang not found")


def build(verbose: bool = False) -> str:
    """Compile the two Fortran sources and link sac_vac (always rebuilds)."""
    os.makedirs(OUT, exist_ok=True)
    fc = flang()
    objs = []
    for src in SOURCES:
        obj = os.path.join(OUT, src[:-4] + ".o")
        cmd = [fc, "-O2", "-ffree-form", "-J", OUT, "-c", os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed for {src}:\n{r.stderr}")
        objs.append(obj)
    if not os.path.exists(os.path.join(LIBDIR, "libsac_amd.so")):
        raise RuntimeError("sac_amd/libsac_amd.so is missing: build the HIP library first")
    cmd = [fc, "-o", EXE] + objs + [f"-L{LIBDIR}", "-lsac_amd", f"-Wl,-rpath,{LIBDIR}",
                                    "-Wl,-rpath,$ORIGIN/../../sac_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    return EXE


if __name__ == "__main__":
    print(build(verbose=True))
