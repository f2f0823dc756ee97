"""Build recipe of the oracle's C++ restatement (TEST AND BASELINE
INFRASTRUCTURE ONLY): oracle/cpu_mpm.cpp -> oracle/libcpu_mpm.so with g++,
OpenMP, float32 without contraction (the cell-key rule of the oracle is
x * inv_dx - 0.5 rounded twice)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_mpm.cpp")
LIB = os.path.join(HERE, "libcpu_mpm.so")


def build(force=True):
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    tmp = LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp",
           "-fPIC", "-shared", SRC, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
