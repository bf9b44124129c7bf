"""Builds libswd.so (HIP kernels for gfx950 + C-ABI) and the AFor / AVSPRTM / Decon drivers in-tree.

    python simple-wavenumber-deconvolution_amd/build.py [--jobs N]

Outputs: simple-wavenumber-deconvolution_amd/lib/libswd.so,
         simple-wavenumber-deconvolution_amd/bin/AFor, .../bin/AVSPRTM, .../bin/Decon
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
DRV = os.path.join(PKG, "drivers")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SWD_ARCH", "gfx950")

# -ffp-contract=off: only the fmas written in the kernels are fused, so the
# arithmetic matches the host oracle operation for operation.  Packed FP32
# instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", f"-I{INC}", f"-I{CSRC}"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp",
             "-Wall", f"-I{INC}", f"-I{CSRC}"]

# extra hipcc flags for experiments (e.g. -DSWD_NT=1)
HIP_FLAGS += os.environ.get("SWD_EXTRA_HIPFLAGS", "").split()

HIP_SOURCES = ["swd_plan.hip", "swd_forward.hip", "swd_rtm.hip", "swd_rtm_updown.hip", "swd_born.hip", "swd_multi.hip", "swd_noalloc.hip",
               "swd_kernels.hip", "swd_loops.hip", "swd_fstep.hip", "swd_fstep_fwd.hip", "swd_fstep_pair.hip",
               "swd_fstep_pair_cig.hip", "swd_fstep_born.hip", "swd_fstep2.hip", "swd_decon.hip", "swd_updown.hip", "swd_cig.hip",
               "swd_adjoint.hip"]
CXX_SOURCES = ["swd_setup.cpp"]
DRIVERS = {"AFor": ["afor.cpp", "driver_common.cpp"], "AVSPRTM": ["avsprtm.cpp", "driver_common.cpp"],
           "Decon": ["decon.cpp", "driver_common.cpp"]}


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit(f"build failed: {cmd[-1]}")
    return r


def build(jobs=8, drivers=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        cmds.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(run, cmds))
    so = os.path.join(LIB, "libswd.so")
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", so + ".tmp", *objs, "-lgomp"])
    os.replace(so + ".tmp", so)
    if drivers:
        for exe, srcs in DRIVERS.items():
            paths = [os.path.join(DRV, s) for s in srcs]
            if all(os.path.exists(p) for p in paths):
                run(["g++", "-O2", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", f"-I{INC}",
                     *paths, "-o", os.path.join(BIN, exe), f"-L{LIB}", "-lswd", "-Wl,-rpath,$ORIGIN/../lib"])
    return so


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-drivers", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, not a.no_drivers))
