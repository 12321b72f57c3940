"""Build the HIP shared library of the decode path (in-tree, gfx950).

    python build.py            # builds lib/libcuphy_hip.so

Every HIP source is compiled with -ffp-contract=off (each float op one
IEEE RN op, as in the oracle) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(OUT_DIR, "libcuphy_hip.so")
BENCH = os.path.join(OUT_DIR, "decode_bench")  # native C-ABI driver (csrc/decode_bench.cpp)
ARCH = os.environ.get("NRLDPC_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "ldpc_decode.hip",
    "ldpc_encode.hip",
    "derate_match.hip",
    "dl_rate_match.hip",
    "crc_decode.hip",
    "pusch_fused.hip",
    "cuphy_api.cpp",
    "harq_server.cpp",
    "tb_params.cpp",
]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_ONLY = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = ["hipcc", "-x", "hip"] + HIP_ONLY + COMMON + ["-c", path, "-o", obj]
    else:
        cmd = ["hipcc", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"]  # roctx ranges (csrc/trace.hpp)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # the driver holds one device kernel of its own (the VALU peak probe)
    cmd = ["hipcc", "-x", "hip"] + HIP_ONLY + ["-O2", "-std=c++17",
           "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "decode_bench.cpp"), "-o", BENCH + ".tmp",
           "-L" + OUT_DIR, "-lcuphy_hip", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("decode_bench build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(BENCH + ".tmp", BENCH)
    if verbose:
        print("built", LIB, "and", BENCH)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
