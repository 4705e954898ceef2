"""Build liblpm.so (gfx950) in-tree with hipcc.

    python build.py            # compile csrc/*.hip in parallel -> liblpm.so next to this file

The translation units are the C ABI (lpm_capi.hip) and the K u launchers with their tile kernels (ku*.hip); objects
go to _build/ and are linked into one shared library.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblpm.so")
SRC_DIR = os.path.join(HERE, "csrc")
OBJ_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
         # packed FP32 instructions are switched off for every kernel of this library
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _msgs(res):
    return [ln for ln in (res.stderr or "").splitlines() if "packed-fp32-ops" not in ln]


def _compile(src, verbose):
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, "\n".join(_msgs(res)[-60:])))
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(SRC_DIR, "*.hip")))
    # the largest units first
    srcs.sort(key=lambda p: (0 if os.path.basename(p).startswith("ku3") else 1, p))
    jobs = jobs or max(1, min(len(srcs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + "\n".join(_msgs(res)[-60:]))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
