"""Build the MI355X fitness library and the drop-in Python extension in place.

    python rils-rols_amd/build.py            # librr_hip.so + rils_rols_cpp*.so
    python rils-rols_amd/build.py --hip      # only the HIP library

Outputs (next to this file):
    librr_hip.so                           C ABI of include/rr_fitness.h (hipcc, gfx950)
    rils_rols_cpp.cpython-*.so             pybind11 module `rils_rols_cpp` (g++, links librr_hip.so)
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["rr_capi.hip", "rr_kernels_fp64.hip", "rr_kernels_fp32.hip"]
HOST_SOURCES = ["tree.cpp", "search.cpp", "serve.cpp", "pybind_module.cpp"]

# -ffp-contract=off: no fused multiply-add unless written as fma(); the FP64
# path must produce the same bits as host code compiled the same way.
# -packed-fp32-ops off: packed FP32 VALU forms are not used on this target.
# -disable-promote-alloca-to-vector: the interpreters keep operand-stack slots
# 0-1 in registers and the rarely used deeper slots in private memory; left
# to itself the compiler turns the deep slots into registers too and the
# kernels lose half their occupancy.
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-disable-promote-alloca-to-vector", "-Wall", "-Wno-unused-function"]
HOST_FLAGS = ["-O3", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-Wall",
              "-Wno-unused-function"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    out = os.path.join(HERE, "librr_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    run([HIPCC] + HIP_FLAGS + ["-shared", "-o", out + ".tmp"] + srcs)
    os.replace(out + ".tmp", out)
    return out


def build_host():
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "rils_rols_cpp" + ext)
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if not srcs:
        return None
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], os.path.join(ROOT, "include")]
    cmd = ["g++"] + HOST_FLAGS + ["-shared", "-o", out + ".tmp"] + srcs
    cmd += ["-I" + i for i in inc]
    cmd += ["-L" + HERE, "-lrr_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    run(cmd)
    os.replace(out + ".tmp", out)
    return out


def main(argv):
    build_hip()
    if "--hip" not in argv:
        build_host()


if __name__ == "__main__":
    main(sys.argv[1:])
