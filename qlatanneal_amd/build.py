"""Build the native parts of qlatanneal_amd in-tree.

    python -m qlatanneal_amd.build          (or __graft_entry__.build())

1. libqanneal_hip.so — HIP kernels + C-ABI (include/qanneal_hip.h), hipcc for gfx950.
2. _qanneal*.so      — the C++17 host classes (include/qanneal/*.hpp) and their pybind11 module,
                       g++, linked against libqanneal_hip.so through an $ORIGIN rpath.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIP_LIB = os.path.join(PKG, "libqanneal_hip.so")
ARCH = os.environ.get("QA_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms are kept out of every kernel (hipcc's host pass warns it does not
    # know the feature; that warning is harmless)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(jobs: int = 8) -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = ["hipcc", *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", s, "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            for p in procs:
                if p.wait() != 0:
                    raise subprocess.CalledProcessError(p.returncode, "hipcc")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, "hipcc")
    _run(["hipcc", *HIPCC_FLAGS, "-shared", *objs, "-o", HIP_LIB])
    return HIP_LIB


def build_module() -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "_qanneal" + suffix)
    srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    if not srcs:
        return ""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           *srcs, "-o", out, f"-L{PKG}", "-lqanneal_hip", "-Wl,-rpath,$ORIGIN",
           f"-Wl,-rpath,{rocm}/lib"]
    _run(cmd)
    return out


def build_all() -> None:
    build_hip()
    build_module()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
