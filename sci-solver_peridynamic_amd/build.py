"""Build the native libraries of the package in-tree.

  libpdsolver.so  HIP kernels + the C ABI of include/pd_solver.h (gfx950)
  libpdhost.so    C++ SystemParameter / Scene / Simulator / DataIO classes
                  (host code above the C ABI, no device code)
  peri_run        command-line driver (reference examples/main.cpp)

Usage: python build.py [--jobs N]
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PD_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # bit-exact parity with the host oracle: no FMA contraction
    "-ffp-contract=off",
    # IEEE float denormals in the float inner solve, as on the host
    "-fno-gpu-flush-denormals-to-zero",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}",
]
HIP_SOURCES = ["pd_step.hip", "pd_bonds.hip", "pd_mixed.hip", "pd_learned.hip", "pd_api.hip"]
HOST_SOURCES = ["parameters.cpp", "scene.cpp", "mesh_fill.cpp", "simulator.cpp", "data_io.cpp",
                "host_capi.cpp"]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{HOST}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({os.path.basename(cmd[-1])})")
    return r


def build(jobs: int = 4, verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    lib = os.path.join(HERE, "libpdsolver.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib + ".tmp", *objs])
    os.replace(lib + ".tmp", lib)
    host_srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if host_srcs:
        hlib = os.path.join(HERE, "libpdhost.so")
        _run([CXX, *CXX_FLAGS, "-shared", "-o", hlib + ".tmp", *host_srcs,
              f"-L{HERE}", "-lpdsolver", "-Wl,-rpath,$ORIGIN"])
        os.replace(hlib + ".tmp", hlib)
        main = os.path.join(HOST, "main.cpp")
        if os.path.exists(main):
            exe = os.path.join(HERE, "peri_run")
            _run([CXX, *CXX_FLAGS, "-o", exe + ".tmp", main, f"-L{HERE}", "-lpdhost", "-lpdsolver",
                  "-Wl,-rpath,$ORIGIN"])
            os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", lib)


if __name__ == "__main__":
    jobs = 4
    if "--jobs" in sys.argv:
        jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(jobs, verbose=True)
