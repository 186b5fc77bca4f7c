"""Build libbecgpp_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m becgpp_amd._build``.  Objects go to ``becgpp_amd/_build/``
(git-ignored); the shared library to ``becgpp_amd/libbecgpp_hip.so``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbecgpp_hip.so")
SOURCES = ["bec_ctrl.hip", "bec_plan.cpp", "bec_solver.cpp", "bec_prof.cpp", "bec_noalloc.hip", "bec_tf.hip"]
# bec_kernels.hip is compiled once per translation unit (its BEC_TU parts: the
# elementwise kernels, one unit per kinetic-pass axis, two for the line passes)
KERNEL_TUS = 8
ARCH = os.environ.get("BECGPP_ARCH", "gfx950")

FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
         # packed FP32 instructions are disabled for all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(job):
    src, tu = job
    stem = os.path.splitext(src)[0] + ("" if tu is None else f"_tu{tu}")
    obj = os.path.join(OBJ, stem + ".o")
    defs = [] if tu is None else [f"-DBEC_TU={tu}"]
    cmd = [_hipcc(), *FLAGS, *defs, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = [("bec_kernels.hip", k) for k in range(KERNEL_TUS)] + [(src, None) for src in SOURCES]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 8))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
