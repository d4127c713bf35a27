"""Build recipe for libconvgemm.so (gfx950 HIP kernels + C host layer).

    python -m convgemm_amd.build            # build in place
    python -m convgemm_amd.build --verbose

Kernels (csrc/kernels/*.hip) are compiled by hipcc for gfx950 only; the host
layer (csrc/host/*.c) is plain C compiled by gcc against the HIP runtime
API; hipcc links both into convgemm_amd/libconvgemm.so.  Objects go to
convgemm_amd/_build/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libconvgemm.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: with AGPR accumulators the loop-carried
    # values were copied AGPR<->VGPR every K step (measured 8x more VALU than
    # MFMA instructions in the BF16 kernels) and occupancy dropped
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result",
]
CC_FLAGS = [
    "-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)
    return out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every kernel and host source and link libconvgemm.so; returns its path."""
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    cc = shutil.which("gcc") or "cc"
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "kernels", "*.hip")))
    c_srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.c")))
    cmds = []
    objs = []
    for s in hip_srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", s, "-o", o])
        objs.append(o)
    for s in c_srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        cmds.append([cc, *CC_FLAGS, "-c", s, "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-ldl"], verbose)
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    print(build(verbose=a.verbose, jobs=a.j))


if __name__ == "__main__":
    sys.exit(main())
