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
                                 "-o", tmp] + srcs + ["-lrocsolver", "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FDRIVER = os.path.join(FORTRAN, "bs_driver")


def build_fortran():
    """The iso_c_binding module and the bs_driver Fortran host (amdflang);
    skipped when no Fortran compiler is installed."""
    fc = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        return None
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    steps = [
        [fc, "-O2", "-c", os.path.join(FORTRAN, "bigstick_amd.f90"), "-o",
         os.path.join(bdir, "bigstick_amd.o"), "-module-dir", bdir],
        [fc, "-O2", "-I", bdir, os.path.join(FORTRAN, "bs_driver.f90"),
         os.path.join(bdir, "bigstick_amd.o"), "-L", HERE, "-lbigstick_amd",
         "-Wl,-rpath,$ORIGIN/..", "-L/opt/rocm/lib", "-lamdhip64", "-o", FDRIVER + ".tmp"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed:\n" + r.stdout + r.stderr)
    os.replace(FDRIVER + ".tmp", FDRIVER)
    return FDRIVER


if __name__ == "__main__":
    print(build_native(verbose=True))
    print(build_fortran())
