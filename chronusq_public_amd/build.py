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


# the sources of the ERI + digestion launches of a Fock build: the kernels,
# their headers and the Rys table
KERNEL_SOURCES = ("eri_", "digest", "kernels.", "device_types.hpp", "rys.hpp", "rys_table.bin", "cand.hip",
                  "limits.hpp")


def source_digest():
    """First 16 hex digits of a SHA-256 over the sources of the ERI and
    digestion kernels (file names included).  Kernel profiles record it,
    so that a profiled figure is reported only beside the kernels it was
    measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in sorted(os.listdir(CSRC)):
        if not f.startswith(KERNEL_SOURCES) or f == "rys_table_data.inc":
            continue
        h.update(f.encode())
        with open(os.path.join(CSRC, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


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


HDF5_PREFIX = os.environ.get("CQX_HDF5_PREFIX", "/opt/conda")
H5LIB = os.path.join(HERE, "libcqx_h5.so")


def build_h5():
    """libcqx_h5.so: the ChronusQ binary-file (HDF5) interop, host C++ linked
    against the HDF5 C library under HDF5_PREFIX; a library of its own so
    that libcqx.so never depends on HDF5.  Returns None (and builds
    nothing) where the HDF5 headers and library are absent."""
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(lib, "libhdf5.so"))):
        print(f"HDF5 not found under {HDF5_PREFIX}: libcqx_h5.so not built", file=sys.stderr)
        return None
    src = os.path.join(HERE, "h5io", "cqx_h5.cpp")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", INCLUDE, "-I", inc, src, "-o", H5LIB + ".tmp",
           "-L", lib, "-lhdf5", "-Wl,-rpath," + lib]
    subprocess.run(cmd, check=True)
    os.replace(H5LIB + ".tmp", H5LIB)
    return H5LIB


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_driver()
    build_h5()
