"""Builds liblbfem_hip.so for gfx950 in-tree with hipcc (no JIT, no cache).

    python lbm-fem_amd/build.py            # -> lbm-fem_amd/lbfem_amd/_lib/liblbfem_hip.so
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lbfem_amd", "_lib")
LIB = os.path.join(OUT_DIR, "liblbfem_hip.so")
SOURCES = ["api.hip", "solver_kernels.hip", "quad_kernels.hip"] + [f"tile_{t}_p{p}.hip" for t in ("f64", "f32") for p in (1, 2, 3)]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    # packed FP32 arithmetic disabled for all device code of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fused only within one source expression: every kernel that evaluates the same
    # expressions rounds them the same way (all tile shapes, the row-march kernel and the batched
    # kernels give bitwise identical populations), whatever code surrounds them
    "-ffp-contract=on",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def build(verbose=True, jobs=None, incremental=False):
    """Compiles every source (incremental=True keeps the objects in _lib/obj and recompiles a source only
    when it, a header or this file is newer than its object)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj") if incremental else OUT_DIR
    os.makedirs(obj_dir, exist_ok=True)
    deps = [os.path.abspath(__file__)] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")] \
        + [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    newest_dep = max(os.path.getmtime(f) for f in deps)
    objs = []
    procs = []
    t0 = time.time()
    for src in SOURCES:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        objs.append(obj)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(
                newest_dep, os.path.getmtime(os.path.join(CSRC, src))):
            continue
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 feature; that warning is expected
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed = True
            print(f"[build] {src} failed:\n" + "\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print(f"[build] {src}:\n" + "\n".join(lines[:40]), file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if not incremental:
        for o in objs:
            os.remove(o)
    # the C++ driver over the C ABI (host code only: g++, no HIP headers)
    drv = os.path.join(OUT_DIR, "cgdbe_hip")
    subprocess.check_call([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-Wextra",
                           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "apps", "cgdbe_hip.cpp"),
                           "-L", OUT_DIR, "-llbfem_hip", "-Wl,-rpath,$ORIGIN", "-o", drv + ".tmp"])
    os.replace(drv + ".tmp", drv)
    if verbose:
        print(f"[build] {LIB} ({time.time() - t0:.1f} s)")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
