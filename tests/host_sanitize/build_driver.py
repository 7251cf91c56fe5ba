"""Build tests/host_sanitize/_bin/abi_driver: the library's four sources with
their host code under AddressSanitizer and UndefinedBehaviorSanitizer (the
device code is compiled as for the library), linked with abi_driver.cpp.
Test infrastructure, run on the CPU: ``python tests/host_sanitize/build_driver.py``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "espressopp_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
BIN = os.path.join(HERE, "_bin")
DRIVER = os.path.join(BIN, "abi_driver")
ROCM = "/opt/rocm"

SOURCES = ["kernels.hip", "sim.hip", "batch.hip", "mlp.hip"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-g", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + INCLUDE,
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
HOST_FLAGS = ["-O1", "-g", "-std=c++17", "-I" + INCLUDE, "-fno-omit-frame-pointer",
              "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


def run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=False):
    os.makedirs(BIN, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(BIN, os.path.splitext(src)[0] + ".o")
        run([os.path.join(ROCM, "bin", "hipcc")] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    clang = os.path.join(ROCM, "llvm", "bin", "clang++")
    drv = os.path.join(BIN, "abi_driver.o")
    run([clang] + HOST_FLAGS + ["-c", os.path.join(HERE, "abi_driver.cpp"), "-o", drv], verbose)
    run([clang] + HOST_FLAGS + objs + [drv, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
                                       "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", DRIVER], verbose)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
