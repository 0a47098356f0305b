"""Compile the occxc library (liboccxc.so) for gfx950 with hipcc, in-tree.

Every HIP source is compiled for gfx950 only; packed-FP32 instruction
selection is disabled for all device code.
"""
import os
import pathlib
import subprocess
import sys
import concurrent.futures as cf

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "liboccxc.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "contract.hip", "grad.hip", "setup.hip", "nlc.hip", "api.cpp",
           "server.cpp"]
COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
          "-Wno-unused-function", "-Wno-unused-command-line-argument",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"-I{HERE.parent / 'include'}"]


def _compile(src):
    obj = BUILD / (src.replace(".", "_") + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=True):
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for (obj, warn), src in zip(results, SOURCES):
        if warn and verbose:
            print(f"[{src}] {warn[-3000:]}", file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp)] + \
          [str(o) for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
