"""Build the native library (libstc.so) and the `stc` command for gfx950.

Every HIP source is compiled by hipcc with --offload-arch=gfx950 and packed
FP32 instructions disabled; host-only sources are compiled by g++.  Outputs
land beside this file (in-tree), so they travel with the source tree.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_SOURCES = ["csr_build.hip", "count.hip", "mlp.hip"]
# measurement helper for bench.py only, built into its own library
PROBE_SOURCE = "bench_probe.hip"
HIP_HOST_SOURCES = ["api.cpp", "upload.cpp"]  # host code that calls the HIP runtime
HOST_SOURCES = ["gen.cpp", "graph_io.cpp"]

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall",
              "-march=x86-64-v2"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src in HIP_SOURCES or src == PROBE_SOURCE:
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", out]
    elif src in HIP_HOST_SOURCES:
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", path, "-o", out]
    else:
        cmd = ["g++"] + HOST_FLAGS + ["-c", path, "-o", out]
    _run(cmd)
    return out


def build(verbose=False):
    """Compile everything; return the path of libstc.so."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at " + HIPCC)
    os.makedirs(OBJ, exist_ok=True)
    srcs = HIP_SOURCES + HIP_HOST_SOURCES + HOST_SOURCES
    with ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(_compile, srcs + [PROBE_SOURCE]))
    probe_obj = objs.pop()
    lib = os.path.join(HERE, "libstc.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
         ["-pthread", "-ldl"])
    os.replace(tmp, lib)
    probe = os.path.join(HERE, "libstc_bench_probe.so")
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", probe + ".tmp", probe_obj])
    os.replace(probe + ".tmp", probe)
    cli_obj = os.path.join(OBJ, "cli.o")
    _run(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, "cli.cpp"), "-o", cli_obj])
    exe = os.path.join(HERE, "stc")
    _run(["g++", "-o", exe + ".tmp", cli_obj, "-L" + HERE, "-lstc",
          "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(exe + ".tmp", exe)
    if verbose:
        print("built", lib, "and", exe)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
