"""Build libs2t.so (HIP, gfx950) and the Node N-API addon in-tree.

Usage: python spectral-2d-turbulence_amd/build.py [--jobs J] [--no-node]

Outputs
  spectral-2d-turbulence_amd/lib/libs2t.so      the C ABI of include/s2t.h
  spectral-2d-turbulence_amd/js/s2t_napi.node   N-API addon over libs2t.so (rpath $ORIGIN/../lib)
"""

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
SOURCES = ["step_kernels.hip", "aux_kernels.hip", "core.hip", "api.hip", "batch.hip",
           "server.hip"]
ARCH = os.environ.get("S2T_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# The step kernels fuse a multiply and an add into an FMA only within one expression: with fusion
# across statements the compiler's choice depends on how often a product is used in the whole
# kernel, so kernels that share the same per-value code (the column-group widths, the
# field-parallel column kernel) could round differently; they give the same state bit for bit.
FILE_FLAGS = {"step_kernels.hip": ["-ffp-contract=on"]}
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), proc.stdout))
    return proc.stdout


def build_lib(jobs=4, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)

    def one(src):
        obj = os.path.join(OBJ, src + ".o")
        out = _run([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
                    "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(one, SOURCES))
    target = os.path.join(LIB, "libs2t.so")
    tmp = target + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp])
    os.replace(tmp, target)
    return target


def build_node(verbose=False):
    """N-API addon: plain g++ against the Node headers, no node-gyp."""
    inc = None
    for cand in ("/usr/include/node", "/usr/local/include/node"):
        if os.path.exists(os.path.join(cand, "node_api.h")):
            inc = cand
            break
    if inc is None:
        if verbose:
            print("node headers not found; skipping the N-API addon")
        return None
    src = os.path.join(PKG, "js", "s2t_napi.cc")
    target = os.path.join(PKG, "js", "s2t_napi.node")
    tmp = target + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I" + inc,
          "-I" + os.path.join(ROOT, "include"), src, "-o", tmp,
          "-L" + LIB, "-ls2t", "-Wl,-rpath,$ORIGIN/../lib", "-pthread"])
    os.replace(tmp, target)
    return target


def build(jobs=4, node=True, verbose=False):
    lib = build_lib(jobs=jobs, verbose=verbose)
    addon = build_node(verbose=verbose) if node else None
    return lib, addon


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--no-node", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args(argv)
    lib, addon = build(jobs=args.jobs, node=not args.no_node, verbose=args.verbose)
    print("built", lib)
    if addon:
        print("built", addon)
    return 0


if __name__ == "__main__":
    sys.exit(main())
