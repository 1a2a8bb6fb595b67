"""Build libcqx.so (HIP for gfx950) in-tree.

Sources: csrc/*.cpp (host orchestration, basis handling, one-electron
integrals, the SCF driver) and csrc/*.hip (device kernels); linked against
rocBLAS and rocSOLVER from the ROCm install.  The Rys table (csrc/rys_table.bin, written by
csrc/gen_rys_table.py) is embedded as a C array.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcqx.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def write_rys_inc():
    tab = np.fromfile(os.path.join(CSRC, "rys_table.bin"), dtype="<f8")
    path = os.path.join(CSRC, "rys_table_data.inc")
    with open(path, "w") as f:
        for i in range(0, tab.size, 4):
            f.write(", ".join(float(v).hex() for v in tab[i:i + 4]) + ",\n")
    return path


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel, then link."""
    write_rys_inc()
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".cpp", ".hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [HIPCC] + compile_flags + ["-I", CSRC, "-I", INCLUDE, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        return obj

    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "lib")
    cmd = ([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs +
           # rocBLAS / rocSOLVER: GEMM and eigensolves of the device-resident SCF
           ["-L", rocm_lib, "-lrocsolver", "-lrocblas", "-Wl,-rpath," + rocm_lib])
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_cpp_driver():
    """The C++ adapter example/test driver (tests/cpp/formgd), linked
    against the in-tree libcqx.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "formgd.cpp")
    out = os.path.join(root, "tests", "cpp", "formgd")
    cmd = ["g++", "-std=c++17", "-O2", "-I", INCLUDE, src, "-o", out, "-L", HERE, "-lcqx",
           "-Wl,-rpath,$ORIGIN/../../chronusq_public_amd"]
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_driver()
