"""Build recipe of the native test programs (tests/native/*.cpp), in-tree under tests/native/_build.

drop_in_driver_{f64,f32}: a reference-shaped C++ caller of the class API, compiled against the
build's headers and linked with librrtmgp_cpp_{f64,f32}.so (run by tests/test_gpu_drop_in.py).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "rte-rrtmgp-cpp_amd")
LIB = os.path.join(PKG, "lib")
OUT = os.path.join(HERE, "_build")


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    outs = []
    for prec, defs in (("f64", []), ("f32", ["-DRTE_USE_SP"])):
        exe = os.path.join(OUT, f"drop_in_driver_{prec}")
        cmd = ["g++", "-std=c++17", "-O1", "-DRTE_USE_CBOOL", *defs, "-I", os.path.join(PKG, "include"),
               "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "drop_in_driver.cpp"), "-o", exe + ".tmp",
               "-L", LIB, f"-lrrtmgp_cpp_{prec}", "-lrrtmgp_hip", f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN/../../../rte-rrtmgp-cpp_amd/lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"native test build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
