"""Build the native library libmaos_tomo.so in-tree (host C++ with g++, HIP
kernels with hipcc for gfx950) and, as test infrastructure, the FP64 CPU oracle.

    python -m maos_amd.build            # both
    python -m maos_amd.build --no-oracle

Outputs: maos_amd/_build/libmaos_tomo.so, oracle/_build/libtomo_oracle.so
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "maos_amd")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(BUILD, "libmaos_tomo.so")

HOST_SRC = ["csrc/host/setup.cpp", "csrc/host/geom.cpp", "csrc/host/fdpcg_setup.cpp", "csrc/host/binio.cpp",
            "csrc/host/fit.cpp", "csrc/host/conf.cpp"]
HIP_SRC = ["csrc/hip/kernels.hip", "csrc/hip/plan.hip", "csrc/hip/fit.hip", "csrc/hip/solve64.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code (wrong results
# have been observed from some of their forms on this hardware).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in HOST_SRC:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(PKG, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SRC:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(PKG, src), "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "_build", "libtomo_oracle.so")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.v))
    if not a.no_oracle:
        print(build_oracle())


if __name__ == "__main__":
    sys.exit(main())
