"""CPU baseline of the direct J/K build: the FP64 oracle timed on the host.

TEST INFRASTRUCTURE ONLY: started as a child process by bench.py's
cpu_baseline leg, never by the product package.  The parent sets
OMP_NUM_THREADS, OMP_PROC_BIND=close and OMP_PLACES=cores in this
process's environment before the OpenMP runtime starts, so the binding
applies to the oracle alone and not to the benchmark's GPU host threads.

usage: python -m oracle.cpu_baseline <in.npz> <out.npz>

in.npz: gbs (basis text), Z, xyz (bohr), mats (stacked NB x NB densities),
types, hers, threads, repeats.  Protocol: one untimed warm-up build, then
`repeats` timed builds; out.npz receives the last build's outputs and
every build time, and stdout one JSON line with the median.
"""
import json
import sys
import time

import numpy as np


def main(argv):
    src, dst = argv[1], argv[2]
    from oracle.oracle import OracleBasis
    d = np.load(src, allow_pickle=False)
    gbs = str(d["gbs"])
    mats = [np.asfortranarray(m) for m in d["mats"]]
    types = [int(t) for t in d["types"]]
    hers = [bool(h) for h in d["hers"]]
    threads, repeats = int(d["threads"]), int(d["repeats"])
    ob = OracleBasis(gbs, d["Z"], d["xyz"])
    ob.prepare()  # Schwarz bounds, once per geometry (untimed, as on the GPU)
    out = ob.contract(mats, types, hers, nthreads=threads)  # warm-up
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        out = ob.contract(mats, types, hers, nthreads=threads)
        times.append(time.perf_counter() - t0)
    np.savez(dst, out=np.stack(out), times=np.array(times))
    print(json.dumps({"median_s": float(np.median(times)), "times_s": times}), flush=True)


if __name__ == "__main__":
    main(sys.argv)
