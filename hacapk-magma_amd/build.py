"""Build script for the H-matvec engine (no cmake: plain compiler invocations).

Products (all in-tree, git-ignored):
  hacapk-magma_amd/lib/libhmv.so     HIP engine for gfx950 (C ABI: include/hmv.h)
  hacapk-magma_amd/lib/libhmvgen.so  CPU input generator   (C ABI: include/hmvgen.h)
  oracle/liboracle.so                CPU oracle (test infrastructure)
  oracle/_ref/libhacapk_ref.so       reference entry-function sources compiled with
                                     flang (only where /root/reference exists)
  oracle/_ref/ref_layout_*.txt       flang's layout of the reference's container types
  oracle/_ref/ref_driver             Fortran caller of libhmv.so with those types

Usage: python hacapk-magma_amd/build.py [--only hmv|hmvgen|oracle|ref]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
INC = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HMV_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC, "-I" + CSRC, "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off"]

HMV_SOURCES = ["hmv_plan.cpp", "hmv_api.cpp", "hmv_file.cpp", "hmv_kernels.hip", "hmv_matmat.hip", "hmv_solver.hip",
               "hmv_block_solver.hip", "hmv_gcr.hip", "hmv_pipe.hip", "hmv_construct.hip",
               "hmv_kstream.hip"]


def run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (p.returncode, " ".join(cmd), out))
    return out


def build_hmv():
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    objs = []
    jobs = []
    for src in HMV_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd.insert(1, "-xhip")
        jobs.append(cmd)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(run, jobs))
    out = os.path.join(LIB, "libhmv.so")
    tmp = out + ".tmp"
    run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs +
        ["-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    return out


def build_hmvgen():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libhmvgen.so")
    tmp = out + ".tmp"
    run(["g++"] + HOST_FLAGS + ["-fopenmp", "-shared", "-I" + INC,
         os.path.join(CSRC, "hgen", "hgen.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    out = os.path.join(ORACLE, "liboracle.so")
    tmp = out + ".tmp"
    run(["gcc", "-O3", "-fPIC", "-std=c11", "-march=x86-64-v3", "-ffp-contract=off",
         "-fopenmp", "-shared", os.path.join(ORACLE, "hmv_oracle.c"), "-o", tmp, "-lm"])
    os.replace(tmp, out)
    return out


def build_ref():
    script = os.path.join(ORACLE, "ref_build.sh")
    if not os.path.isdir("/root/reference"):
        return None
    run(["bash", script])
    return os.path.join(ORACLE, "_ref", "libhacapk_ref.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["hmv", "hmvgen", "oracle", "ref"], default=None)
    args = ap.parse_args(argv)
    # the reference step links oracle/_ref/ref_driver against libhmv.so: engine first
    steps = {"hmvgen": build_hmvgen, "oracle": build_oracle, "hmv": build_hmv, "ref": build_ref}
    for name, fn in steps.items():
        if args.only and args.only != name:
            continue
        out = fn()
        print("built", name, "->", out, flush=True)


if __name__ == "__main__":
    main()
