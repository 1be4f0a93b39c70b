"""Build libgsbn_amd.so in-tree: the gfx950 kernels with hipcc and the host
framework with g++, both with floating-point contraction disabled so the host
and device arithmetic round identically."""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgsbn_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("GSBN_OFFLOAD_ARCH", "gfx950")

FP_FLAGS = ["-ffp-contract=off", "-fno-fast-math"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + FP_FLAGS
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include"), "-Wall", "-Wno-unused-function"] + FP_FLAGS

HIP_SOURCES = ["kernels.hip", "recall.hip", "noalloc.hip"]
HOST_SOURCES = ["framework.cpp", "procedures.cpp", "proc_upd_lazy.cpp", "capi.cpp", "coeff.cpp",
                "snapshot.cpp", "structural_plasticity.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o], o))
    cxx = shutil.which("g++") or "g++"
    for s in HOST_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(([cxx] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o], o))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + [o for _, o in jobs]
         + ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
