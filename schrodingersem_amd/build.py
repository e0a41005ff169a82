"""Build libsem_hip.so (gfx950) in-tree with hipcc, and the CPU oracle.

Usage: python -m schrodingersem_amd.build [--jobs N]
The shared library lands next to this file; nothing is installed.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsem_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for all device code. a*b+c is
# contracted into an FMA only within one source expression (-ffp-contract=on):
# every compiled form of a kernel (work shapes, step modes, batched or single
# instances, the Krylov recurrence in whichever kernel evaluates it) then does
# the same arithmetic and gives the same bits.
COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-result",
]

# (source, object, extra flags)
UNITS = [
    ("kernels_op.hip", "kernels_op_f64.o", []),
    ("kernels_op.hip", "kernels_op_f32.o", ["-DSEM_OP_FP32"]),
    ("kernels_vec.hip", "kernels_vec.o", []),
    ("engine.hip", "engine.o", []),
    ("mesh_host.cpp", "mesh_host.o", ["-x", "hip"]),
    ("mlp.hip", "mlp.o", []),
]


def _deps_mtime():
    """newest modification time of the headers every unit may include"""
    t = 0.0
    for d in (CSRC, os.path.join(ROOT, "include")):
        for f in os.listdir(d):
            if f.endswith((".hpp", ".h")):
                t = max(t, os.path.getmtime(os.path.join(d, f)))
    return t


def _compile(unit, force=True):
    src, obj, extra = unit
    out = os.path.join(OBJ, obj)
    srcp = os.path.join(CSRC, src)
    if not force and os.path.exists(out) and os.path.getmtime(out) > max(os.path.getmtime(srcp), _deps_mtime()):
        return out
    cmd = [HIPCC] + COMMON + extra + ["-c", srcp, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return out


def build_lib(jobs=None, force=True):
    """Compile every unit (force) or only those older than their source or a header."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(UNITS), os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _compile(u, force), UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    return LIB


def build_cpp_tests():
    """The C++ drop-in test (reference test sequence against include/sem.hpp)."""
    out_dir = os.path.join(ROOT, "tests", "cpp", "bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "test_drop_in")
    cmd = ["g++", "-std=c++20", "-O2", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "test_drop_in.cpp"), "-o", out, "-L" + HERE, "-lsem_hip",
           "-Wl,-rpath,$ORIGIN/../../../schrodingersem_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C++ drop-in test build failed:\n%s" % r.stderr[-4000:])
    return out


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s" % r.stderr[-4000:])
    return os.path.join(ROOT, "oracle", "liboracle_sem.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--skip-oracle", action="store_true")
    ap.add_argument("--incremental", action="store_true", help="recompile only out-of-date units")
    a = ap.parse_args(argv)
    print(build_lib(a.jobs, force=not a.incremental))
    print(build_cpp_tests())
    if not a.skip_oracle:
        print(build_oracle())


if __name__ == "__main__":
    sys.exit(main())
