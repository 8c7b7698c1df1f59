"""Build recipe for libprr.so (hipcc, gfx950) and the CPU oracle.

``python -m photorealisticrender4610_amd.build`` compiles every source under
csrc/ in parallel into _lib/libprr.so. Device code is built with
-ffp-contract=off (no FMA contraction: geometry is bit-identical to the
reference's CPU arithmetic) and with packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libprr.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PRR_OFFLOAD_ARCH", "gfx950")
COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def sources():
    out = []
    for dp, _, fns in os.walk(CSRC):
        for fn in sorted(fns):
            if fn.endswith((".hip", ".cpp")):
                out.append(os.path.join(dp, fn))
    return sorted(out)


def _compile(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    obj = os.path.join(BUILD, rel + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build_lib(jobs=None, verbose=True):
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_dropin_test(verbose=True):
    """tests/cpp/dropin_main.cpp: a scene script against the C++ drop-in API
    (include/prr/prr.hpp), linked to libprr.so."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
    exe = os.path.join(OUT, "dropin_main")
    cmd = ["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), src, "-o", exe,
           "-L" + OUT, "-lprr", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in test build failed: %s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", exe)
    return exe


def build_oracle(verbose=True):
    mk = shutil.which("make")
    cmd = [mk, "-C", os.path.join(ROOT, "oracle"), "-j4", "oracle"]
    if os.path.isdir("/root/reference"):
        cmd.append("ref")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (r.stdout, r.stderr))
    if verbose:
        print("built oracle")


if __name__ == "__main__":
    build_lib()
    build_dropin_test()
    if "--no-oracle" not in sys.argv:
        build_oracle()
