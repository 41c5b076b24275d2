"""Build recipe of the gfx950 libraries (in-tree, no JIT cache).

  lib/librrtmgp_hip.so      HIP kernels + the thin C ABI of include/rrtmgp_hip.h (f32 and f64)
  lib/librrtmgp_cpp_f64.so  C++ class API (Gas_optics_rrtmgp_gpu, ...) built with Float=double,
                            plus the C facade of include/rrtmgp_api.h
  lib/librrtmgp_cpp_f32.so  the same with Float=float (RTE_USE_SP)

All HIP code is compiled with the packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SRC = os.path.join(HERE, "src")
INC_ROOT = os.path.join(ROOT, "include")
INC_CPP = os.path.join(HERE, "include")
LIB_DIR = os.path.join(HERE, "lib")
BUILD_DIR = os.path.join(HERE, "lib", "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INC_ROOT, "-I", CSRC, "-Wno-unused-result"]

KERNEL_SOURCES = ["runtime.hip", "array_kernels.hip", "gas_optics_kernels.hip", "optical_props_kernels.hip",
                  "rte_solver_kernels.hip", "rte_lw_scattering_kernels.hip", "rte_adjoint_kernels.hip",
                  "mlp_kernels.hip"]
CPP_SOURCES = []   # filled when the class layer sources exist


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def lib_path(name):
    return os.path.join(LIB_DIR, name)


def build_kernels(verbose=False, jobs=8):
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    cmds = []
    for src in KERNEL_SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    out = lib_path("librrtmgp_hip.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_cpp(verbose=False, jobs=8):
    srcs = sorted(f for f in os.listdir(SRC) if f.endswith(".cpp")) if os.path.isdir(SRC) else []
    if not srcs:
        return []
    outs = []
    for prec, defs in (("f64", []), ("f32", ["-DRTE_USE_SP"])):
        objs, cmds = [], []
        for s in srcs:
            obj = os.path.join(BUILD_DIR, f"{s[:-4]}_{prec}.o")
            objs.append(obj)
            cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-DRTE_USE_CBOOL", *defs,
                         "-I", INC_ROOT, "-I", INC_CPP, "-I", "/opt/rocm/include",
                         "-D__HIP_PLATFORM_AMD__", "-c", os.path.join(SRC, s), "-o", obj])
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        out = lib_path(f"librrtmgp_cpp_{prec}.so")
        _run(["g++", "-shared", "-fPIC", *objs, "-o", out + ".tmp", "-L", LIB_DIR, "-lrrtmgp_hip",
              "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64"], verbose)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    k = build_kernels(verbose)
    c = build_cpp(verbose)
    return [k, *c]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
