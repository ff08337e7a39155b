"""Build libmpm_amd.so (the HIP path and its C ABI) in-tree for gfx950.

Run as ``python -m code_mpm_amd.build`` or through ``__graft_entry__.build()``.
Always recompiles.  Flags:
  -ffp-contract=off   no FMA fusion: every FP64 operation rounds once, as the
                      reference's x86-64 build does;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      keeps packed FP32 instructions out of the code object.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmpm_amd.so")

SOURCES = ["mpm_kernels.hip", "mpm_engine.hip", "mpm_mlp.hip", "mpm_surrogate.hip", "mpm_eso.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmpm_amd.so")
    os.replace(tmp, LIB)
    return LIB


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libmpm_physics.so")
BIN = os.path.join(HERE, "bin")
GXX_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def build_host(verbose=False):
    """The C++ host classes (libmpm_physics.so, over libmpm_amd.so) and the
    reference-style bar driver (bin/mpm_bar_demo), with g++: the host side
    holds no HIP types, only the C ABI."""
    inc = ["-I", os.path.join(os.path.dirname(HERE), "include")]
    rocm_lib = "/opt/rocm/lib"
    cmds = [
        ["g++"] + GXX_FLAGS + ["-fPIC", "-shared", os.path.join(HOST, "mpm_physics.cpp")] + inc +
        ["-L", HERE, "-lmpm_amd", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + rocm_lib, "-o", HOST_LIB + ".tmp"],
    ]
    os.makedirs(BIN, exist_ok=True)
    demo = os.path.join(BIN, "mpm_bar_demo")
    cmds.append(["g++"] + GXX_FLAGS + [os.path.join(HOST, "bar_demo.cpp")] + inc +
                ["-L", HERE, "-lmpm_physics", "-lmpm_amd", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link," + rocm_lib,
                 "-o", demo + ".tmp"])
    for k, cmd in enumerate(cmds):
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("g++ failed building the C++ host side")
        if k == 0:
            os.replace(HOST_LIB + ".tmp", HOST_LIB)
    os.replace(demo + ".tmp", demo)
    return HOST_LIB, demo


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_host(verbose=True))
