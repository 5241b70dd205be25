"""Build recipe of the oracle (test infrastructure): compiles wf_oracle.c into
oracle/build/libwf_oracle.so with gcc (FP64, OpenMP, no FMA contraction).

The reference itself cannot be compiled here: its CUDA path needs nvcc and
cuFFT, and its C/OpenMP paths need FFTW3 and no longer compile against the
current headers, so no oracle/_ref build exists.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "build")
LIB = os.path.join(OUT_DIR, "libwf_oracle.so")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=gnu11",
           "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-variable",
           os.path.join(HERE, "wf_oracle.c"), "-o", LIB + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
