"""Builds la3_amd/libla3.so for gfx950 in-tree.

host_build.cpp (ingest, R-MAT) is plain C++ built with g++; capi.hip (C ABI,
kernels) is built with hipcc for gfx950 only. Both link into one shared
library with a C ABI (include/la3.h).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libla3.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # Packed FP32 instructions stay off: plain float code only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wextra"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    host_o = os.path.join(BUILD, "host_build.o")
    serve_o = os.path.join(BUILD, "server.o")
    capi_o = os.path.join(BUILD, "capi.o")
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "host_build.cpp"), "-o", host_o])
    _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "server.cpp"), "-o", serve_o])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "capi.hip"), "-o", capi_o])
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", capi_o, host_o, serve_o, "-o", tmp,
          "-pthread", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, OUT)
    # C++ drivers over include/la3.hpp (the reference's app mains).
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "apps", "la3_run.cpp"),
          "-o", os.path.join(HERE, "bin", "la3_run"), f"-L{HERE}", "-lla3",
          "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    # R-MAT generator as a tool (writes LA3 binary edge files).
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "apps", "la3_rmat.cpp"),
          "-o", os.path.join(HERE, "bin", "la3_rmat"), f"-L{HERE}", "-lla3",
          "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    # Format tools (the reference's mtx2bin / bin2mtx), host only.
    for tool in ("la3_mtx2bin", "la3_bin2mtx"):
        _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "apps", tool + ".cpp"),
              "-o", os.path.join(HERE, "bin", tool)])
    return OUT


def build_program(src, out):
    """Compiles a source file of VertexProgram<P> programs
    (include/la3_vertex_program.hpp) into a shared library linked against
    libla3.so: the way a caller builds programs of its own for the engine."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    rel = os.path.relpath(HERE, os.path.dirname(os.path.abspath(out)))
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{inc}", src, "-o", tmp, f"-L{HERE}", "-lla3",
          f"-Wl,-rpath,$ORIGIN/{rel}", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
