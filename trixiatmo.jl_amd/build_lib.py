"""Build libatmo_hip.so for gfx950 in place (no cmake; hipcc + g++ only).

Usage: python build_lib.py [--jobs J]
The library is written next to this file; objects go to ./_build/.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libatmo_hip.so")
# the HDF5 solution/restart file library (host only); it links the HDF5 C
# library of the image's conda tree and is skipped where that is absent
H5_LIB = os.path.join(HERE, "libatmo_h5.so")
H5_SRC = os.path.join(CSRC, "h5", "atmo_h5.cpp")
H5_PREFIX = os.environ.get("ATMO_HDF5_PREFIX", "/opt/conda")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]
# per compile unit: the FP64 3D stage kernels scheduled for instruction-level
# parallelism (the same results, the same 4 waves per SIMD; on MI355X over
# four alternating pairs of runs 0.3-0.6 % less time per polydeg-4 shell
# stage and about 1 % at polydeg 7)
HIP_UNIT_FLAGS = {
    "euler3d_fp64.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    steps = []
    objs = []
    for name in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, name)
        base, ext = os.path.splitext(name)
        if ext == ".hip":
            obj = os.path.join(BUILD, base + ".hip.o")
            steps.append([HIPCC, *HIP_FLAGS, *HIP_UNIT_FLAGS.get(name, []), "-c", src, "-o", obj])
        elif ext == ".cpp":
            obj = os.path.join(BUILD, base + ".cpp.o")
            steps.append(["g++", *CXX_FLAGS, "-c", src, "-o", obj])
        else:
            continue
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for r in ex.map(_run, steps):
            if verbose and (r.stdout or r.stderr):
                sys.stdout.write(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    build_h5()
    return LIB


def build_h5():
    """libatmo_h5.so; returns its path, or None when no HDF5 library is found"""
    inc = os.path.join(H5_PREFIX, "include")
    libdir = os.path.join(H5_PREFIX, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and
            os.path.exists(os.path.join(libdir, "libhdf5.so"))):
        sys.stderr.write(f"HDF5 not found under {H5_PREFIX}: libatmo_h5.so is not built\n")
        return None
    tmp = H5_LIB + ".tmp"
    # the C++ runtime is linked in statically: the only library taken from the
    # conda tree is libhdf5 (and what it needs), never a second libstdc++
    _run(["g++", *CXX_FLAGS, "-shared", f"-I{inc}", H5_SRC, "-o", tmp,
          "-static-libstdc++", "-static-libgcc",
          os.path.join(libdir, "libhdf5.so"), f"-Wl,-rpath,{libdir}", "-Wl,--enable-new-dtags"])
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
