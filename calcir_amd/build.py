"""Build libcalcir.so (HIP, gfx950) and the calcir CLI in-tree with hipcc.

Every HIP source is compiled for gfx950 with packed-FP32 instructions
disabled; host sources are compiled by the same driver as plain C++.
Run:  python -m calcir_amd.build   (or via __graft_entry__.build()).
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcalcir.so")
CLI = os.path.join(HERE, "bin", "calcir")

HIP_SOURCES = ["k_sites.hip", "k_gemm.hip", "k_eigh.hip", "k_band.hip", "k_spectra.hip", "pipeline.hip",
               "adjoint.hip", "noalloc.hip"]
CXX_SOURCES = ["common.cpp", "tuner.cpp", "host_io.cpp", "synth.cpp", "xtc.cpp", "cli.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-Wall", "-Wno-unused-result"]
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
HIP_FLAGS = ["--offload-arch=gfx950", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + HIP_FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + HOST_FLAGS + ["-x", "c++", "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    srcs = HIP_SOURCES + CXX_SOURCES
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    cli_obj = os.path.join(BUILD, "calcir_cli_main.o")
    _run([HIPCC] + COMMON + HOST_FLAGS + ["-x", "c++", "-c", os.path.join(CSRC, "calcir_cli_main.cpp"), "-o", cli_obj])
    _run([HIPCC, "-o", CLI + ".tmp", cli_obj, "-L" + HERE, "-lcalcir", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(CLI + ".tmp", CLI)
    if verbose:
        print("built", LIB, "and", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
