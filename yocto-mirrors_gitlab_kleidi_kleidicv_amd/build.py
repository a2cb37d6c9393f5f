"""Build libkleidicv_hip.so (gfx950) in-tree with hipcc.

    python yocto-mirrors_gitlab_kleidi_kleidicv_amd/build.py

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled and floating-point contraction off (the float gathers
fuse exactly where the reference's AArch64 build does, through explicit
fmaf calls).  Objects go to _build/ next to this file; the shared library is
written beside this file so it ships with the tree.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkleidicv_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["filters.hip", "resize.hip", "transform.hip", "yuv.hip", "batch.hip",
           "prepare.cpp", "api.cpp", "batch_api.cpp", "serve.cpp", "hal.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    # 8- and 16-byte loads and stores through 1-byte-aligned vector types are
    # deliberate: gfx950 vector memory runs in unaligned mode, and rows of
    # any stride are read that way
    "-Wno-align-mismatch",
    "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, proc


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    sources = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        results = list(pool.map(_compile, sources))
    failed = False
    for src, _, proc in results:
        noise = [l for l in proc.stderr.splitlines()
                 if "packed-fp32-ops" not in l and l.strip()]
        if proc.returncode != 0:
            failed = True
            sys.stderr.write("hipcc failed for %s:\n%s\n" % (src, proc.stderr))
        elif verbose and noise:
            sys.stderr.write("\n".join(noise) + "\n")
    if failed:
        raise RuntimeError("HIP build failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + \
        [obj for _, obj, _ in results]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
