"""Build libfemhip.so (HIP kernels + C-ABI) in-tree for gfx950.

``python -m fempar_amd.build`` or ``fempar_amd.build.build()``.  hipcc
cross-compiles without a GPU, so this runs in a CPU-only container too.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfemhip.so")
SOURCES = ["api.cpp", "space.cpp", "setup.hip", "sliced.cpp", "assemble.hip", "cg.hip", "batch.hip", "serve.cpp", "norms.hip", "stencil.hip", "noalloc.hip", "multifield.hip", "mapped.hip"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _flags():
    return [
        "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
        # keep packed-FP32 instructions out of generated code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
        "-Wall", "-Wno-unused-result", "-Wno-unused-variable",
    ]


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    cmds = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmds.append([hipcc, *_flags(), *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)

    def run(cmd):
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("compile failed: " + " ".join(cmd) + "\n" + p.stdout + p.stderr)
        if verbose and (p.stdout or p.stderr):
            sys.stderr.write(p.stdout + p.stderr)

    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(run, cmds))
    tmp = LIB + ".tmp"
    run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_EXE = os.path.join(HERE, "test_poisson_hip")


SEAMS_EXE = os.path.join(HERE, "test_poisson_seams")


def build_fortran() -> str:
    """Compile the ISO_C_BINDING module, the FEMPAR plug-in seams with their
    HIP subtypes, and the two test_poisson drivers with AMD flang."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    obj = os.path.join(mod, "femhip_names.o")
    seams = os.path.join(mod, "fempar_seams.o")
    plugins = os.path.join(mod, "hip_fempar_plugins.o")
    link = ["-J", mod, "-L" + HERE, "-lfemhip", "-Wl,-rpath,$ORIGIN"]
    for cmd in ([fc, "-O2", "-c", os.path.join(FORTRAN_DIR, "femhip_names.f90"), "-J", mod, "-o", obj],
                [fc, "-O2", os.path.join(FORTRAN_DIR, "test_poisson_hip.f90"), obj, *link, "-o", FORTRAN_EXE],
                [fc, "-O2", "-c", os.path.join(FORTRAN_DIR, "fempar_seams.f90"), "-J", mod, "-o", seams],
                [fc, "-O2", "-c", os.path.join(FORTRAN_DIR, "hip_fempar_plugins.f90"), "-J", mod, "-o", plugins],
                [fc, "-O2", os.path.join(FORTRAN_DIR, "test_poisson_seams.f90"), obj, seams, plugins, *link,
                 "-o", SEAMS_EXE]):
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("Fortran build failed: " + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    return FORTRAN_EXE


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran())
