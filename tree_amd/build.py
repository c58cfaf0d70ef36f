"""Compile libtreemi.so (HIP for gfx950) in-tree.

Every .hip file under csrc/ is compiled by hipcc for gfx950 only; host-only
.cpp files by g++.  Objects go to _build/ at the repository root and the
shared library to tree_amd/libtreemi.so, next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtreemi.so")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off: some of their forms have given wrong
    # results on this hardware without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include")]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    else:
        cmd = ["g++"] + CXX_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_fortran(verbose)
    return OUT


FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FSRC = ["datastr.f90", "treemi_capi.f90", "treemi_force.f90", "main.f90"]
FBIN = os.path.join(HERE, "bin", "octree_gpu")


def build_fortran(verbose=False):
    """The Fortran driver (tree_amd/bin/octree_gpu), linked to libtreemi.so."""
    if not os.path.exists(FLANG):
        if verbose:
            print("amdflang not found: Fortran driver skipped")
        return None
    fdir = os.path.join(BUILD, "fortran")
    os.makedirs(fdir, exist_ok=True)
    os.makedirs(os.path.dirname(FBIN), exist_ok=True)
    objs = []
    for f in FSRC:
        src = os.path.join(HERE, "fortran", f)
        obj = os.path.join(fdir, f.replace(".f90", ".o"))
        r = subprocess.run([FLANG, "-O2", "-c", src, "-o", obj, "-module-dir", fdir], capture_output=True,
                           text=True, cwd=fdir)
        if r.returncode != 0:
            raise RuntimeError("Fortran compile failed: %s\n%s" % (f, r.stderr))
        objs.append(obj)
    tmp = FBIN + ".tmp"
    r = subprocess.run([FLANG, "-o", tmp] + objs + ["-L" + HERE, "-ltreemi", "-Wl,-rpath,$ORIGIN/.."],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran link failed:\n%s" % r.stderr)
    os.replace(tmp, FBIN)
    if verbose:
        print("built", FBIN)
    return FBIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
