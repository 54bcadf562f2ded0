"""Build the oracle's compiled CPU restatement (test infrastructure only).

oracle/cpu_step.cpp -> oracle/_build/pecs_cpu_step, linked against the host
setup objects of the native build (mesh, local operators, supernodal
Cholesky), which pecs_amd.build compiles first.  Only tests/ and the
cpu_baseline leg of bench.py run it.

    python -m oracle.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
EXE = os.path.join(BUILD, "pecs_cpu_step")
HOST_OBJ = [os.path.join(ROOT, "pecs_amd", "_build", o) for o in ("mesh.o", "ops.o", "cholesky.o")]


def build(verbose=False):
    for o in HOST_OBJ:
        if not os.path.exists(o):
            raise RuntimeError(f"{o} missing: build the native library first (pecs_amd.build)")
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-mavx2", "-mfma", os.path.join(HERE, "cpu_step.cpp")] + \
        HOST_OBJ + ["-o", EXE]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    return EXE


if __name__ == "__main__":
    print(build(verbose=True))
