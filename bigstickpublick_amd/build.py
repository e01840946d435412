"""Builds the native library in-tree: bigstickpublick_amd/libbigstick_amd.so
(hipcc, gfx950) and the Fortran iso_c_binding module when amdflang exists."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbigstick_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off (see the engine's build notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-Wall", "-Wno-unused-result",
]


def build_native(verbose=False):
    srcs = [os.path.join(CSRC, "bs_engine.hip"), os.path.join(CSRC, "bs_tables.cpp")]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(os.path.dirname(HERE), "include"),
                                 "-o", tmp] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_native(verbose=True))
