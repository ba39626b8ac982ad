"""In-tree build of libpcx_gpu.so for gfx950 (hipcc cross-compiles; no GPU needed).

The kernels are instantiated in several translation units (csrc/k_*.hip) that compile in
parallel; pcx_gpu.hip holds the C ABI.  Objects go to packed_amd/_build/; with force=False a
unit is recompiled only when it or a header it may include is newer than its object.
"""
from __future__ import annotations

import glob
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libpcx_gpu.so")

# the slowest units first
SOURCES = ["k_mfma_big.hip", "k_mfma_p1.hip", "k_mfma_p4.hip", "k_f64.hip", "k_shapes_alt.hip", "k_mfma.hip", "k_f64_alt.hip",
           "k_shapes.hip", "k_single.hip", "k_xc.hip", "k_misc.hip", "pcx_gpu.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _stale(src: str, obj: str, deps_mtime: float) -> bool:
    if not os.path.exists(obj):
        return True
    m = os.path.getmtime(obj)
    return m < os.path.getmtime(src) or m < deps_mtime


def _compile(src: str, obj: str) -> None:
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}", "-c", src, "-o", obj + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n" + res.stdout + res.stderr)
    os.replace(obj + ".tmp", obj)


def build(verbose: bool = False, force: bool = True, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + [os.path.join(ROOT, "include", "pcx_gpu.h")]
    deps_mtime = max(os.path.getmtime(h) for h in headers)
    todo, objs = [], []
    for name in SOURCES:
        src = os.path.join(CSRC, name)
        obj = os.path.join(OBJ, name.replace(".hip", ".o"))
        objs.append(obj)
        if force or _stale(src, obj, deps_mtime):
            todo.append((src, obj))
    jobs = jobs or max(1, min(len(todo), os.cpu_count() or 1, 8))
    if todo:
        with ThreadPoolExecutor(jobs) as ex:
            for fut in [ex.submit(_compile, s, o) for s, o in todo]:
                fut.result()
    if todo or not os.path.exists(OUT):
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", OUT + ".tmp"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("link failed:\n" + res.stdout + res.stderr)
        os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(f"built {len(todo)} of {len(SOURCES)} units -> {OUT}")
    return OUT


if __name__ == "__main__":
    import sys

    print(build(verbose=True, force="--force" in sys.argv))
