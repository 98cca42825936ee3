"""In-tree build of the native libraries (gfx950).

  libsayaka_dev.so   hipcc, every .hip kernel for gfx950, FP64 parity kernels with
                     -ffp-contract=off (no FMA contraction), packed-FP32 codegen disabled,
                     small per-thread arrays kept in registers (PROMOTE_FLAGS)
  libsayaka_host.so  g++, host tree / tables / solver driver (C ABI include/sayaka_host.h),
                     linked against libsayaka_dev.so
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

DEV_SOURCES = ["dev/sayaka_dev.hip"]
DEV_HEADERS = ["dev/dev_common.h", "dev/kernels.h", "dev/bottom.h", "dev/coef.h", "dev/mixed.h"]
HOST_SOURCES = ["host/amr_tree.cpp", "host/capi_tree.cpp", "host/tree_data.cpp", "host/mac_solver.cpp",
                "host/capi_mac.cpp", "host/h5lite.cpp", "host/writer_hdf5.cpp", "host/fill_patch.cpp"]
DRIVERS = {"sayaka_c1_solve": "drivers/c1_solve.cpp", "sayaka_projection_step": "drivers/projection_step.cpp"}


# Small per-thread index arrays that are indexed by a run-time axis (the halo gathers' i/j/k
# triples) are promoted by LLVM to LDS by default.  That LDS copy is addressed by a flat thread
# id built from the workgroup sizes, which the kernel reads from the dispatch packet: one extra
# scalar load per wave, measured at ~14 us per stencil launch on MI355X (a 300-block launch took
# 21.6 us, 7.6 us without the halo gather).  Promotion to registers instead costs no VGPRs here.
PROMOTE_FLAGS = ["-mllvm", "-disable-promote-alloca-to-lds", "-mllvm", "-amdgpu-promote-alloca-to-vector-limit=1024"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")
    return r


def _dev_cmd(out, defines=()):
    return [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *PROMOTE_FLAGS, *defines, f"-I{INC}",
            *[os.path.join(CSRC, s) for s in DEV_SOURCES], "-o", out + ".tmp"]


def build_dev(verbose=False, phases=True) -> str:
    """libsayaka_dev.so and, beside it (phases), libsayaka_dev_phases.so: the same library with
    the coarsest-level solve's phase timing compiled in (include/sayaka_phases.h), built
    concurrently."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libsayaka_dev.so")
    outp = os.path.join(LIB, "libsayaka_dev_phases.so")
    procs = [(out, subprocess.Popen(_dev_cmd(out), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))]
    if phases:
        procs.append((outp, subprocess.Popen(_dev_cmd(outp, ["-DSAYAKA_BOTTOM_PHASES"]), stdout=subprocess.PIPE,
                                             stderr=subprocess.PIPE, text=True)))
    failed = []
    for o, p in procs:
        so, se = p.communicate()
        if verbose:
            print(" ".join(p.args), flush=True)
        if p.returncode != 0:
            sys.stderr.write(so + se)
            failed.append(o)
    if failed:
        raise RuntimeError(f"build failed: {failed}")
    for o, _ in procs:
        os.replace(o + ".tmp", o)
    return out


def build_host(verbose=False) -> str:
    """Host library; links libsayaka_dev.so (the MacSolver drives the device through its C ABI)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libsayaka_host.so")
    cmd = ["g++", "-std=c++17", "-O2", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{INC}",
           f"-I{os.path.join(INC, 'sayaka')}", *[os.path.join(CSRC, s) for s in HOST_SOURCES],
           f"-L{LIB}", "-lsayaka_dev", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    _run(cmd, verbose)
    os.replace(out + ".tmp", out)
    return out


def build_drivers(verbose=False):
    """C++ drivers written against the MacSolver / AmrTree / TreeData headers."""
    outs = []
    for name, src in DRIVERS.items():
        out = os.path.join(LIB, name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{INC}", f"-I{os.path.join(INC, 'sayaka')}",
               os.path.join(CSRC, src), f"-L{LIB}", "-lsayaka_host", "-lsayaka_dev", "-Wl,-rpath,$ORIGIN",
               "-o", out + ".tmp"]
        _run(cmd, verbose)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_all(verbose=False):
    return [build_dev(verbose), build_host(verbose), *build_drivers(verbose)]


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print("built", p)
