#!/usr/bin/env python3
"""Direct Fock-build benchmark on one MI355X.

Workload (default): GHF/X2C two-component direct J/K build in FP64 --
(H2O)_16 placed from seed 0 in a 12 A cube (O-O >= 2.6 A), cc-pVDZ
(384 spatial functions, spinor dimension 768); complex Hermitian density
D = C C^H with C 768 x 160, real and imaginary parts N(0, 1/sqrt(768));
its Pauli components S, Z, Y, X (4 x 384^2 complex).  One step = one
formGD: the five contractions J[S], K[S], K[Z], K[Y], K[X] of
GTODirectTPIContraction::twoBodyContract with Schwarz threshold 1e-12,
densities and outputs resident in HBM.

Prints one JSON line.  Roofline: the ERI/digestion kernels' algorithmic
FP64 flops (the Rys flop model over the quartets evaluated) divided by
their device time from HIP events on the launch stream.  CPU baseline: the
FP64 oracle (a C/OpenMP restatement of the reference scaffold) timed on one
full build of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Fock builds/sec, mixed vs FP64 at |ΔE|≤1µEh; SCF p50/p99 latency"
FP64_VECTOR_PEAK = 78.6  # TFLOP/s, MI355X


def pauli_density(nb, nocc, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n2 = 2 * nb
    C = rng.normal(0, 1 / np.sqrt(n2), size=(n2, nocc)) + 1j * rng.normal(0, 1 / np.sqrt(n2), size=(n2, nocc))
    D = C @ C.conj().T
    aa, ab, ba, bb = D[:nb, :nb], D[:nb, nb:], D[nb:, :nb], D[nb:, nb:]
    # SpinScatter (src/cqlinalg/blasutil.cxx:71-189): S, Z, Y, X
    return [aa + bb, aa - bb, 1j * (ab - ba), ab + ba]


def workload(name):
    from chronusq_public_amd.chem import basis_text, water_cluster
    if name == "ghf16":
        mol = water_cluster(16, 12.0, seed=0)
        return mol, basis_text("cc-pVDZ"), "2c"
    if name == "rhf4":
        mol = water_cluster(4, 8.0, seed=0)
        return mol, basis_text("cc-pVDZ"), "rhf"
    raise SystemExit(f"unknown workload {name}")


def densities(kind, nb):
    if kind == "2c":
        S, Z, Y, X = pauli_density(nb, 160, 0)
        mats = [S, S, Z, Y, X]
        types = [0, 1, 1, 1, 1]
        return mats, types, True
    rng = np.random.Generator(np.random.PCG64(0))
    C = rng.normal(0, 1 / np.sqrt(nb), size=(nb, 20))
    D = C @ C.T
    D = 0.5 * (D + D.T)
    return [D, D], [0, 1], False


def cpu_baseline(gbs, mol, mats, types, cplx, gpu_out):
    from oracle.oracle import OracleBasis
    threads = min(16, os.cpu_count() or 1)
    ob = OracleBasis(gbs, mol.Z, mol.xyz)
    ob.prepare()  # Schwarz, once per geometry (untimed, as on the GPU side)
    t0 = time.perf_counter()
    ref = ob.contract(mats, types, [True] * len(mats), nthreads=threads)
    dt = time.perf_counter() - t0
    err = max(float(np.abs(g - r).max()) for g, r in zip(gpu_out, ref))
    return {"value": 1.0 / dt, "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": "one full Fock build of the same workload (FP64 C/OpenMP oracle)",
            "seconds": dt}, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ghf16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM "
                         "counter calibration for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU: the direct J/K engine runs one process "
                         "on one MI355X (multi-GPU sharding is not part of this build)")

    from chronusq_public_amd import cqx
    mol, gbs, kind = workload(args.workload)
    basis = cqx.BasisSet(gbs, mol)
    tpi = cqx.DirectTPI(basis, 1e-12, device=0)
    gto = cqx.GTODirectTPIContraction(tpi)
    nb = basis.nBasis
    mats, types, cplx = densities(kind, nb)
    dX = [cqx.DeviceMatrix(tpi, nb, cplx).upload(m) for m in mats]
    dA = [cqx.DeviceMatrix(tpi, nb, cplx) for _ in mats]
    cont = [cqx.TwoBodyContraction(x, a, True, t) for x, a, t in zip(dX, dA, types)]

    def step():
        for a in dA:
            a.zero()
        gto.twoBodyContract(cont)

    if args.copy_mb:
        import ctypes as C
        nbytes = args.copy_mb << 20
        src, dst = C.c_void_p(), C.c_void_p()
        cqx._check(cqx.lib().cqx_device_alloc(tpi._h, nbytes, C.byref(src)), tpi._h)
        cqx._check(cqx.lib().cqx_device_alloc(tpi._h, nbytes, C.byref(dst)), tpi._h)
        cqx._check(cqx.lib().cqx_memcpy(tpi._h, dst, src, nbytes, 2), tpi._h)
        cqx.lib().cqx_device_free(tpi._h, src)
        cqx.lib().cqx_device_free(tpi._h, dst)
    for _ in range(args.warmup):
        step()
    eri_ms, flops = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        st = tpi.last_stats()
        eri_ms.append(st["ms_eri"])
        flops.append(st["flops_model"])
    dt = time.perf_counter() - t0
    st = tpi.last_stats()
    ms_step = 1e3 * dt / args.steps
    eri_avg = float(np.mean(eri_ms))
    achieved = float(np.mean(flops)) / (eri_avg * 1e-3) / 1e12
    out = [a.download() for a in dA]

    line = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("GHF 2c J/K, (H2O)16 seed 0 in 12 A cube, cc-pVDZ (NB 384, spinor 768), "
                         "5 contractions J[S] K[S] K[Z] K[Y] K[X], Schwarz 1e-12"
                         if args.workload == "ghf16" else
                         "RHF J/K, (H2O)4 seed 0 in 8 A cube, cc-pVDZ (NB 96), 2 contractions"),
            "precision": "fp64",
            "nbasis": nb,
            "quartets": st["n_quartets"],
            "candidates": st["n_candidates"],
        },
        "roofline": {
            "bound": "mfma",
            "compute": "fp64 vector ALU (MI355X fp64 vector and matrix peaks are both 78.6 TFLOP/s)",
            "kernel": "ERI + J/K digestion launches of one build (all angular-momentum classes)",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK,
            "traffic": None,
            "flops_per_build": float(np.mean(flops)),
            "eri_ms_per_build": eri_avg,
        },
    }
    if not args.no_cpu:
        cb, err = cpu_baseline(gbs, mol, mats, types, cplx, out)
        line["cpu_baseline"] = cb
        line["max_abs_dJK_vs_oracle"] = err
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
