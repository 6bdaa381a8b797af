"""Build the C restatement of the randomized step (cpu_edrsvd.c) into oracle/libcpu_edrsvd.so.

TEST INFRASTRUCTURE ONLY (the CPU baseline bench.py times, checked by the tests).  gcc, OpenMP,
-O3 with AVX2/FMA code (no -march=native: the library is built here and runs on the GPU box's host)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_edrsvd.c")
LIB = os.path.join(HERE, "libcpu_edrsvd.so")


def build():
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O3", "-mavx2", "-mfma", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-o", tmp, SRC, "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle CPU build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
