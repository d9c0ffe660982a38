"""Builds libsep.so, the MI355X seplib hot-path library, in-tree.

Device code (csrc/hip/*.hip) is compiled by hipcc for gfx950 only; the C99
host layer (csrc/host/*.c) by gcc, the reference's own toolchain; hipcc
links both into seplib_amd/libsep.so.  Objects go to seplib_amd/_build/.

    python -m seplib_amd.build          # or: python seplib_amd/build.py
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsep.so")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

# -ffp-contract=off: every product and sum is rounded on its own, as in the
# reference's C; the neighbour test r2 < (cf+skin)^2 must see the same bits.
# -packed-fp32-ops off: packed FP32 VALU forms are not trusted on this pool.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
C_FLAGS = ["-std=c99", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
           "-Wall", "-Wextra", "-Wno-unused-parameter", "-D_GNU_SOURCE"]


def _sources():
    hip = sorted(os.path.join(CSRC, "hip", f) for f in os.listdir(os.path.join(CSRC, "hip"))
                 if f.endswith(".hip"))
    c = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
               if f.endswith(".c"))
    return hip, c


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    inc = ["-I" + INCLUDE, "-I" + CSRC]
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + inc + ["-c", src, "-o", obj]
    else:
        cmd = [CC] + C_FLAGS + inc + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=True, jobs=8):
    """Compiles every source and links seplib_amd/libsep.so (always)."""
    os.makedirs(BUILD, exist_ok=True)
    hip, c = _sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, hip + c))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=True):
    """Builds the CPU oracle (and the compiled reference when present)."""
    odir = os.path.join(ROOT, "oracle")
    if shutil.which("make") is None:
        raise RuntimeError("make not found")
    out = _run(["make", "-C", odir, "-j8"])
    if verbose and out.strip():
        print(out.strip().splitlines()[-1])


if __name__ == "__main__":
    build()
    if "--no-oracle" not in sys.argv:
        build_oracle()
