"""Build libocpx.so (HIP kernels for gfx950 + the C++ host code).

Usage: python -m opencaeporox_amd.csrc.build  (or __graft_entry__.build()).
Each translation unit is compiled in parallel with hipcc, then linked into
opencaeporox_amd/lib/libocpx.so.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT_DIR = os.path.join(PKG, "lib")
OBJ_DIR = os.path.join(PKG, "lib", "obj")
SOURCES = ["pe.hip", "pe_team.hip", "pe_sta.hip", "tail.hip", "fim.hip", "solver.hip", "wells.hip", "noalloc.hip", "cpr.hip", "ocpx_host.cpp"]
ARCH = os.environ.get("OCPX_ARCH", "gfx950")


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if p and os.path.exists(p):
            return p
    raise RuntimeError("hipcc not found")


def flags():
    return ["-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
            # packed FP32 instructions are kept off; this code is FP64 anyway
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(PKG, "..", "include")]


def _stale(src, obj):
    """The object is older than its source or any header of this directory."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(HERE, src)] + [os.path.join(HERE, h) for h in os.listdir(HERE) if h.endswith(".hpp")]
    deps.append(os.path.join(PKG, "..", "include", "ocpx.h"))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def compile_one(src, force=True):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    if not force and not _stale(src, obj):
        return obj
    cmd = [hipcc()] + flags() + ["-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False, force=True):
    """Compile every source (force) or only the out-of-date ones, then link."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    workers = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "4")))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda s: compile_one(s, force), SOURCES))
    out = os.path.join(OUT_DIR, "libocpx.so")
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True, force="--changed" not in sys.argv)
    sys.exit(0)
