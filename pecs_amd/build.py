"""Build the native library (libpecs_amd.so) and the native self-test, in-tree.

Host-only setup code (meshes, local operators, nested dissection, multifrontal
Cholesky) is compiled with g++ -fopenmp; device code and the C-ABI with hipcc
for gfx950.  The C++ host classes and the pecs_run driver (pecs_amd/host) link
against the library.  Outputs go to pecs_amd/ (the .so) and pecs_amd/_build/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpecs_amd.so")
SELFTEST = os.path.join(BUILD, "pecs_selftest")
HOST = os.path.join(HERE, "host")
RUNNER = os.path.join(BUILD, "pecs_run")
HOST_TEST = os.path.join(BUILD, "pecs_host_test")

HOST_SRC = ["mesh.cpp", "ops.cpp", "cholesky.cpp", "vtu.cpp", "dealii_layout.cpp"]
HIP_SRC = ["kernels.hip", "devfactor.hip", "mlp.hip", "abi.hip"]
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-mavx2", "-mfma"]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objs = []
    for s in HOST_SRC:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        _run(["g++"] + CXXFLAGS + ["-c", os.path.join(CSRC, s), "-o", o], verbose)
        objs.append(o)
    hobjs = []
    for s in HIP_SRC:
        o = os.path.join(BUILD, s.replace(".hip", ".hip.o"))
        _run([hipcc] + HIPFLAGS + ["-c", os.path.join(CSRC, s), "-o", o], verbose)
        hobjs.append(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-o", tmp] + hobjs + objs +
         ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-lgomp", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    _run(["g++"] + CXXFLAGS + [os.path.join(CSRC, "selftest.cpp")] + objs + ["-o", SELFTEST], verbose)
    # host classes in the reference roles + the command-line driver, linked to the library
    link = ["-L" + HERE, "-lpecs_amd", "-Wl,-rpath,$ORIGIN/.."]
    _run(["g++", "-O2", "-std=c++17", os.path.join(HOST, "main.cpp"), "-o", RUNNER] + link, verbose)
    _run(["g++", "-O2", "-std=c++17", os.path.join(HOST, "host_test.cpp"), "-o", HOST_TEST] + link, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
