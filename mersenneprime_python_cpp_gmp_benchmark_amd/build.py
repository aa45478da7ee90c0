"""In-tree build of the native parts, for gfx950.

  libllhip.so                          HIP kernels + the C ABI of include/llhip.h (hipcc)
  lucaslehmer<EXT_SUFFIX>              pybind11 drop-in for the reference `lucaslehmer`
  lucaslehmer_cuda<EXT_SUFFIX>         pybind11 batch module under the reference's GPU module name

Everything lands next to this file so it travels with the source tree.  Run
``python -m mersenneprime_python_cpp_gmp_benchmark_amd.build`` or call ``build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("LLH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["ll_large_g0.hip", "ll_large_g1.hip", "ll_large_g2.hip", "ll_large_g3.hip", "ll_batch.hip",
               "ll_large.hip", "ll_mlp.hip", "ll_capi.hip", "ll_tables.cpp"]
# packed-FP32 instructions are switched off for every device compile
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I", INCLUDE, "-I", CSRC]

LIB = os.path.join(HERE, "libllhip.so")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _ext(name: str) -> str:
    return os.path.join(HERE, name + sysconfig.get_config_var("EXT_SUFFIX"))


def build_lib(jobs: int = 4) -> None:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS]
        if src.endswith(".cpp"):
            cmd += ["-x", "hip"]
        cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("hipcc failed")
    for pr in procs:
        if pr.wait() != 0:
            raise RuntimeError("hipcc failed")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", LIB, *objs])


def build_modules() -> None:
    import pybind11
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}", f"-I{INCLUDE}"]
    common = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden", *inc]
    link = [f"-L{HERE}", "-lllhip", "-Wl,-rpath,$ORIGIN"]
    _run([*common, os.path.join(CSRC, "lucaslehmer_module.cpp"), "-o", _ext("lucaslehmer"), *link])
    _run([*common, os.path.join(CSRC, "lucaslehmer_cuda_module.cpp"), "-o", _ext("lucaslehmer_cuda"), *link])


def build() -> None:
    build_lib()
    build_modules()


if __name__ == "__main__":
    build()
    sys.exit(0)
