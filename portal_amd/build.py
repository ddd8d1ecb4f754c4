"""Build the gfx950 shared library ``portal_amd/libportal_hip.so`` with hipcc.

Every source under ``portal_amd/csrc`` is compiled for gfx950 only.  The FP64
parity kernels (walk, inward FP64 integration, SE, image) are built with
``-ffp-contract=off`` so that products round as in the reference; the
mixed-precision and MLP kernels may contract into FMA.  Packed-FP32 instruction
selection is disabled for all HIP code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libportal_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

NO_CONTRACT = {"walk.hip", "rt_f64.hip", "se.hip", "image.hip", "lambda.hip", "ctx.cpp", "se_tables.cpp",
               "star_dirs.cpp"}


def _flags(name: str) -> list[str]:
    f = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include"),
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
    f.append("-ffp-contract=off" if name in NO_CONTRACT else "-ffp-contract=fast")
    if name.endswith(".cpp"):
        f = ["-x", "hip"] + f
    return f


def _compile(name: str) -> tuple[str, int, str]:
    src = os.path.join(CSRC, name)
    obj = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC] + _flags(name) + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return obj, p.returncode, (" ".join(cmd) + "\n" + p.stdout + p.stderr) if p.returncode else p.stderr


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(n for n in os.listdir(CSRC) if n.endswith((".hip", ".cpp")))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        res = list(ex.map(_compile, srcs))
    errs = [r for r in res if r[1]]
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(e[2] for e in errs))
    if verbose:
        for r in res:
            if r[2].strip():
                print(r[2])
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [r[0] for r in res]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    build_fortran()
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
DEMO = os.path.join(HERE, "portal_f90_demo")


def build_fortran() -> str | None:
    """The ISO_C_BINDING module and the Fortran demo caller (when a Fortran compiler is present)."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(OBJ, exist_ok=True)
    mod_o = os.path.join(OBJ, "portal_hip_mod.o")
    subprocess.run([FLANG, "-O2", "-fPIC", "-c", os.path.join(FDIR, "portal_hip_mod.f90"), "-J", OBJ, "-o", mod_o],
                   check=True, capture_output=True)
    subprocess.run([FLANG, "-O2", "-I", OBJ, os.path.join(FDIR, "portal_f90_demo.f90"), mod_o, "-L", HERE,
                    "-lportal_hip", "-Wl,-rpath,$ORIGIN", "-o", DEMO], check=True, capture_output=True)
    return DEMO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
