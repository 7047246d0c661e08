"""Build recipe for the in-tree gfx950 library and the drop-in CLI.

    python -m softmold_amd.build        # or __graft_entry__.build()

Compiles softmold_amd/csrc/* with hipcc for gfx950 into softmold_amd/libsoftmold_hip.so
and the host programs softmold_amd/md_hip (the `MD name` replacement), vesicle_radius and
change_tension (the reference's analysis/vesicleRadius and analysis/changeTension).  Packed FP32 instructions are
disabled for all device code (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsoftmold_hip.so")
# host programs linked against the library: source -> executable in softmold_amd/
TOOLS = {"md_hip.cpp": "md_hip", "vesicle_radius.cpp": "vesicle_radius", "change_tension.cpp": "change_tension"}

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
          f"-I{os.path.join(os.path.dirname(HERE), 'include')}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["sm_kernels.hip", "sm_pair.hip", "sm_sort.hip", "sm_adjoint.hip", "sm_noalloc.hip", "sm_capi.cpp", "sm_mpd_api.cpp",
               "mpd_io.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + DEVICE + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    srcs = LIB_SOURCES + [s for s in TOOLS if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [objs[s] for s in LIB_SOURCES]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(tmp, LIB)
    for s, exe in TOOLS.items():
        if s in objs:
            out = os.path.join(HERE, exe)
            cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", out + ".tmp", objs[s], "-L" + HERE, "-lsoftmold_hip",
                   "-Wl,-rpath,$ORIGIN"]
            subprocess.run(cmd, check=True, capture_output=not verbose)
            os.replace(out + ".tmp", out)
    if verbose:
        print("built", LIB)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
