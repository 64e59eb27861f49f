"""Builds, in-tree:
  libclsph_amd/libclsph_hip.so   the HIP kernels + C ABI (hipcc, gfx950)
  libclsph_amd/libclsph.so       the drop-in C++ API (sph_simulation, scene,
                                 houdini_file_saver) over the C ABI (g++)
  libclsph_amd/bin/clsph_particles  the command-line driver (examples/)

    python -m libclsph_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libclsph_hip.so")

SOURCES = ["grid.hip", "sort.hip", "sph.hip", "advect.hip", "nn.hip", "gnn.hip", "capi.cpp",
           "nn_capi.cpp", "gnn_capi.cpp"]
API = os.path.join(HERE, "api")
API_SOURCES = ["sph_simulation.cpp", "scene.cpp", "json.cpp",
               "file_save_delegates/houdini_file_saver.cpp", "util/cl_boilerplate.cpp"]
API_OUT = os.path.join(HERE, "libclsph.so")
CLI_SRC = os.path.join(os.path.dirname(HERE), "examples", "particles.cpp")
CLI_OUT = os.path.join(HERE, "bin", "clsph_particles")
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CLSPH_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-shared", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless.
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError(f"build of {what} failed")
    if verbose and msgs:
        print("\n".join(msgs))


def build(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    _run([HIPCC] + FLAGS + srcs + ["-Wl,-soname,libclsph_hip.so", "-o", tmp],
         "libclsph_hip.so", verbose)
    os.replace(tmp, OUT)

    api = [os.path.join(API, s) for s in API_SOURCES]
    tmp = API_OUT + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra"] + api +
         [OUT, "-Wl,-soname,libclsph.so", "-Wl,-rpath,$ORIGIN", "-pthread", "-o", tmp],
         "libclsph.so", verbose)
    os.replace(tmp, API_OUT)

    os.makedirs(os.path.dirname(CLI_OUT), exist_ok=True)
    tmp = CLI_OUT + ".tmp"
    _run([CXX, "-O2", "-std=c++17", "-Wall", CLI_SRC, API_OUT, OUT,
          "-Wl,-rpath,$ORIGIN/..", "-pthread", "-o", tmp], "clsph_particles", verbose)
    os.replace(tmp, CLI_OUT)
    return OUT


if __name__ == "__main__":
    print(build())
