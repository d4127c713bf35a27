"""Build recipe for libconvgemm.so (gfx950 HIP kernels + C host layer).

    python -m convgemm_amd.build            # build in place
    python -m convgemm_amd.build --verbose

Kernels (csrc/kernels/*.hip) are compiled by hipcc for gfx950 only; the
implicit-GEMM instantiations (csrc/kernels/slice.hip) are compiled once per
(operand type, operation, kernel family) slice so they build in parallel, and
the single-launch batch kernel (csrc/kernels/single.hip) once per operand type.
The host layer (csrc/host/*.c) is plain C compiled by gcc against the HIP
runtime API; hipcc links both into convgemm_amd/libconvgemm.so.  Objects go
to convgemm_amd/_build/.  build_sanitized() makes an ASan/UBSan build of the
host layer for the sanitizer test.
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


# Instantiation slices of the implicit-GEMM kernels (kernels/slice.hip), each
# compiled as its own translation unit so they build in parallel:
# (type name, element type, operation, family 0 = general / 1 = pipelined).
# FP8 serves the forward operations only; the pipelined kernels serve the
# forward pass (NHWC for every type, NCHW for FP32).
SLICE_TYPES = [("f32", "float"), ("bf16", "uint16_t"), ("fp8", "uint8_t")]

# Per operand type extra flags.  FP8: without machine sinking.  The loads the
# pipelined kernels issue with an out-of-range offset for a dead step become
# uniform branches, and the pass then moved the FP8 MFMAs (32-byte operands)
# across them, past the scheduling fences of their read groups, up to the
# end of the loop body: every group's operands were live at once and the
# larger tiles spilled (up to 300 VGPRs to scratch).  Without it no FP8
# pipelined kernel spills; FP32 and BF16 keep their MFMAs in place.
TYPE_FLAGS = {"fp8": ["-mllvm", "-disable-machine-sink"]}


def slices():
    out = []
    for tn, ct in SLICE_TYPES:
        ops = range(2) if tn == "fp8" else range(6)
        for op in ops:
            out.append((tn, ct, op, 0))
        out.append((tn, ct, 0, 1))
        if tn == "f32":
            out.append((tn, ct, 1, 1))
        if tn != "fp8":
            out.append((tn, ct, 2, 1))
    # the pipelined slices are the largest translation units: start them first
    return sorted(out, key=lambda s: (-s[3], s[2]))


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every kernel and host source and link libconvgemm.so; returns its path."""
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    cc = shutil.which("gcc") or "cc"
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    hip_srcs = sorted(s for s in glob.glob(os.path.join(CSRC, "kernels", "*.hip"))
                      if os.path.basename(s) not in ("slice.hip", "single.hip"))
    c_srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.c")))
    cmds = []
    objs = []
    slice_src = os.path.join(CSRC, "kernels", "slice.hip")
    for tn, ct, op, fam in slices():  # the largest translation units first
        o = os.path.join(BUILD, f"slice_{tn}_{op}_{fam}.o")
        cmds.append([hipcc, *HIPCC_FLAGS, *TYPE_FLAGS.get(tn, []), f"-DCG_SLICE_T={ct}", f"-DCG_SLICE_TN={tn}",
                     f"-DCG_SLICE_OP={op}",
                     f"-DCG_SLICE_FAM={fam}", "-c", slice_src, "-o", o])
        objs.append(o)
    single_src = os.path.join(CSRC, "kernels", "single.hip")
    for tn, ct in SLICE_TYPES:  # the single-launch batch kernel, one per operand type
        o = os.path.join(BUILD, f"single_{tn}.o")
        cmds.append([hipcc, *HIPCC_FLAGS, *TYPE_FLAGS.get(tn, []), f"-DCG_SINGLE_T={ct}", f"-DCG_SINGLE_TN={tn}", "-c",
                     single_src, "-o", o])
        objs.append(o)
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


# Host-layer sanitizer build (gcc's AddressSanitizer + UndefinedBehaviorSanitizer,
# host C only: the kernel objects are the normal ones).
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"]


def build_sanitized(out_dir: str, verbose: bool = False) -> str:
    """libconvgemm.so with the host C compiled under ASan + UBSan, in out_dir (with
    the data folder linked beside it, where the library looks for its cost model).
    Reuses the kernel objects of the normal build (building them if absent).  Load it
    with CONVGEMM_LIB=<path> and gcc's libasan preloaded."""
    def kernel_objs():
        return sorted(glob.glob(os.path.join(BUILD, "slice_*.o")) + glob.glob(os.path.join(BUILD, "single_*.o")) +
                      glob.glob(os.path.join(BUILD, "*.hip.o")))
    kobjs = kernel_objs()
    if not kobjs:
        build(verbose=verbose)
        kobjs = kernel_objs()
    cc = shutil.which("gcc") or "cc"
    os.makedirs(out_dir, exist_ok=True)
    objs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "host", "*.c"))):
        o = os.path.join(out_dir, os.path.basename(src) + ".o")
        _run([cc, *CC_FLAGS, *SAN_FLAGS, "-c", src, "-o", o], verbose)
        objs.append(o)
    rt = [subprocess.run([cc, f"-print-file-name={n}"], stdout=subprocess.PIPE, text=True).stdout.strip()
          for n in ("libasan.so", "libubsan.so")]
    lib = os.path.join(out_dir, "libconvgemm.so")
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib, *objs, *kobjs, "-lpthread", "-ldl", *rt],
         verbose)
    data = os.path.join(out_dir, "data")
    if not os.path.exists(data):
        os.symlink(os.path.join(PKG, "data"), data)
    return lib


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    print(build(verbose=a.verbose, jobs=a.j))


if __name__ == "__main__":
    sys.exit(main())
