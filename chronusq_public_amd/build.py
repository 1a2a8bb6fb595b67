"""Build libcqx.so (HIP for gfx950) in-tree.

Sources: csrc/*.cpp (host orchestration and basis handling) and csrc/*.hip
(device kernels).  The Rys table (csrc/rys_table.bin, written by
csrc/gen_rys_table.py) is embedded as a C array.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcqx.so")
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


def build(verbose=False):
    write_rys_inc()
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".cpp", ".hip")))
    cmd = [HIPCC] + FLAGS + ["-I", CSRC, "-o", LIB + ".tmp"] + srcs
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
