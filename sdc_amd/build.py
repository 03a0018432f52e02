"""Builds libsdc_rbc.so (gfx950) in-tree with hipcc.

Usage: python -m sdc_amd.build [--jobs N]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsdc_rbc.so")

SOURCES = ["rbc_setup.cpp", "rbc_abi.cpp", "rbc_io.cpp", "rbc_ensemble.cpp", "rbc_driver.cpp", "rbc_kernels.hip",
           "rbc_velocity.hip", "rbc_closure.hip"]
RUN_EXE = os.path.join(HERE, "rbc_run")
EXTRACT_EXE = os.path.join(HERE, "rbc_extract_vel")
ARCH = os.environ.get("SDC_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
          "-I", CSRC, "-Wall", "-Wno-unused-result",
          # packed FP32 ops stay disabled on this platform
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs=4, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", LIB + ".tmp"] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    # the executables: main()s around rbc_run_main and rbc_extract_vel_main, next to the library
    for src, exe in (("rbc_run.cpp", RUN_EXE), ("rbc_extract_vel.cpp", EXTRACT_EXE)):
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, src),
               "-L", HERE, "-lsdc_rbc", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", LIB, RUN_EXE, EXTRACT_EXE)
    return LIB


if __name__ == "__main__":
    jobs = 4
    if "--jobs" in sys.argv:
        jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(jobs=jobs, verbose=True)
