"""Build the HIP library in-tree: ``libfieldspice_hip.so`` for gfx950.

Run as ``python -m fieldspice_amd.build`` (also called by
``__graft_entry__.build()``).  hipcc cross-compiles for gfx950 without a GPU.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libfieldspice_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["fs_kernels.hip", "fs_transient.hip", "fs_mg.hip", "fs_nn.hip", "fs_pc.hip",
           "fs_defl.hip", "fs_api.hip"]
ARCH = os.environ.get("FIELDSPICE_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the ROCm toolchain is required to build")


OBJ_DIR = os.path.join(HERE, "_build")

FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 instructions are kept out of every kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-result",
         f"-I{INCLUDE}", f"-I{CSRC}"]


def command(out: str = LIB_PATH) -> list[str]:
    """One hipcc line compiling and linking every source (for reference)."""
    return [hipcc(), *FLAGS, "-shared", *[os.path.join(CSRC, s) for s in SOURCES], "-o", out]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stderr[-6000:]}")
    noise = "packed-fp32-ops' is not a recognized feature"
    err = "\n".join(ln for ln in proc.stderr.splitlines() if noise not in ln)
    if verbose and err.strip():
        print(err[-3000:], file=sys.stderr)


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every source to an object in parallel, then link the library
    (every translation unit launches only its own kernels: no relocatable
    device code is needed)."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc(), *FLAGS, "-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1), 8)
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB_PATH + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
