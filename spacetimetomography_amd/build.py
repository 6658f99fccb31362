"""Build recipe of the native library (libstt.so) for gfx950.

Host geometry and synthetic-input code is compiled with g++ and
-ffp-contract=off (the per-view ray geometry must follow the reference's fp32
operation order); the HIP kernels and the C-ABI are compiled with hipcc for
gfx950 only.  Everything lands in-tree under spacetimetomography_amd/lib/.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libstt.so")

ARCH = os.environ.get("STT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", shutil.which("g++") or "g++")

# Device code flags: packed fp32 ops disabled (see the build notes), correctly
# rounded fp32 division/sqrt (ray set-up parity), hardware fp32 atomics.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-Wall"]

HOST_SOURCES = ["stt_geometry.cpp", "stt_synth.cpp", "stt_io.cpp"]
HIP_SOURCES = ["stt_kernels.hip", "stt_capi.cpp"]
# the command-line driver: host code over the C-ABI only
CLI_SOURCE = "stt_cli.cpp"
CLI = os.path.join(LIBDIR, "stt-tomography")


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout)
    return proc.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp"])
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-Wall", "-pthread", "-o", tmp, os.path.join(CSRC, CLI_SOURCE),
          "-L" + LIBDIR, "-lstt", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    if verbose:
        print("built", LIB, "and", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
