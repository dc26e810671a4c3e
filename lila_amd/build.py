"""Build liblila_hip.so (gfx950) in-tree.

HIP sources (*.hip) are compiled with hipcc for gfx950 with packed FP32
disabled; host C sources (*.c) with gcc; everything is linked into
lila_amd/liblila_hip.so with hipcc.  Objects go to lila_amd/_build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblila_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-munsafe-fp-atomics",
]
C_FLAGS = ["-O2", "-std=c11", "-fPIC", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for f in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, f)
        base, ext = os.path.splitext(f)
        obj = os.path.join(BUILD, base + ".o")
        if ext == ".hip":
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, src])
        elif ext == ".c":
            jobs.append(["gcc"] + C_FLAGS + ["-c", "-o", obj, src])
        else:
            continue
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lrocblas"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
