"""Build recipe of the FP64 CPU oracle (test infrastructure).

Compiles oracle/rrtmgp_oracle.cpp into oracle/_build/librrtmgp_oracle.so with
``-O3 -march=x86-64-v3 -ffp-contract=off`` (no FMA contraction, so index arithmetic is
reproducible; x86-64-v3 rather than -march=native because the library is built here and run on
another host).

The reference's own CPU path is not built: its hot arithmetic is Fortran from the un-vendored
rte-rrtmgp submodule and gfortran, boost and netCDF are absent, so there is no oracle/_ref.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "librrtmgp_oracle.so")
SRC = os.path.join(HERE, "rrtmgp_oracle.cpp")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-std=c++17", "-fPIC",
           "-shared", "-fopenmp", SRC, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
