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
their device time from HIP events on the launch stream; HBM traffic,
executed FP64 FLOPs and memory-side atomics come from the committed
counter profile of the same sources, when there is one.  CPU baseline:
the FP64 oracle (a C/OpenMP restatement of the reference scaffold) on the
host's physical cores (OMP_PROC_BIND=close, OMP_PLACES=cores), one warm-up
and the median of five full builds of the same workload.  host_pointer:
the same builds through host matrices (the ChronusQ adapter's call),
PCIe copies included.

The "serving" object reports SCF p50/p99 latency and jobs/s of the serving
loop (see serving()).  The "mixed" object of the line is the
mixed-precision comparison on the
linear alkane C40H82 (all-trans, C-C 1.53 A, C-H 1.09 A, seed 0 +-0.05 A
jitter), 6-31G* (724 functions), RHF density D = C C^T with C 724 x 161,
N(0, 1/sqrt(724)): builds/s of the FP64 build and of the mixed build
(quartets with Schwarz product < 1e-4 in FP32), max|dF| (F - H = G = 2J - K)
and |dE| = 1/4 |Tr(D dG)| of the mixed result against the FP64 one.
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
COULOMB_, EXCHANGE_ = 0, 1
FP64_VECTOR_PEAK = 78.6  # TFLOP/s, MI355X
FP32_VECTOR_PEAK = 157.3  # TFLOP/s, MI355X
LAUNCH_STREAMS = 4  # streams the ERI launches of a build are dealt over (--streams)


def pauli_density(nb, nocc, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n2 = 2 * nb
    C = rng.normal(0, 1 / np.sqrt(n2), size=(n2, nocc)) + 1j * rng.normal(0, 1 / np.sqrt(n2), size=(n2, nocc))
    D = C @ C.conj().T
    aa, ab, ba, bb = D[:nb, :nb], D[:nb, nb:], D[nb:, :nb], D[nb:, nb:]
    # SpinScatter (src/cqlinalg/blasutil.cxx:71-189): S, Z, Y, X
    return [aa + bb, aa - bb, 1j * (ab - ba), ab + ba]


def workload(name):
    from chronusq_public_amd.chem import alkane_chain, basis_text, water_cluster
    if name == "alkane40":
        return alkane_chain(40, 0), basis_text("6-31G*"), "rhf"
    if name == "ghf16":
        mol = water_cluster(16, 12.0, seed=0)
        return mol, basis_text("cc-pVDZ"), "2c"
    if name == "rhf4":
        mol = water_cluster(4, 8.0, seed=0)
        return mol, basis_text("cc-pVDZ"), "rhf"
    raise SystemExit(f"unknown workload {name}")


def densities(kind, nb, nocc=20):
    if kind == "2c":
        S, Z, Y, X = pauli_density(nb, 160, 0)
        mats = [S, S, Z, Y, X]
        types = [0, 1, 1, 1, 1]
        return mats, types, True
    rng = np.random.Generator(np.random.PCG64(0))
    C = rng.normal(0, 1 / np.sqrt(nb), size=(nb, nocc))
    D = C @ C.T
    D = 0.5 * (D + D.T)
    return [D, D], [0, 1], False


def host_cpu():
    """Logical CPUs this process may run on, the physical cores among them
    (distinct package/core ids), and the CPU model name."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return len(cpus), len(cores), model


def cpu_baseline(gbs, mol, mats, types, cplx, gpu_out, repeats=5):
    """The FP64 oracle (a C/OpenMP restatement of the reference scaffold)
    on the host's physical cores: a child process with OMP_PROC_BIND=close
    and OMP_PLACES=cores, one warm-up build, then the median of `repeats`
    builds of the same workload.  Threads: every physical core this process
    may use, or OMP_NUM_THREADS where the environment sets it (the host's
    share of CPUs on a shared machine)."""
    import subprocess
    import tempfile
    nproc, phys, model = host_cpu()
    env_threads = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env_threads) if env_threads.isdigit() and int(env_threads) > 0 else phys
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as tmp:
        src, dst = os.path.join(tmp, "in.npz"), os.path.join(tmp, "out.npz")
        np.savez(src, gbs=np.array(gbs), Z=np.asarray(mol.Z, dtype=np.float64),
                 xyz=np.asarray(mol.xyz, dtype=np.float64), mats=np.stack(mats),
                 types=np.array(types), hers=np.ones(len(mats), dtype=bool),
                 threads=np.array(threads), repeats=np.array(repeats))
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", src, dst], cwd=ROOT, env=env,
                           capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError("cpu baseline failed:\n" + r.stderr[-2000:])
        res = json.loads(r.stdout.strip().splitlines()[-1])
        ref = list(np.load(dst)["out"])
    err = max(float(np.abs(g - x).max()) for g, x in zip(gpu_out, ref))
    return {"value": 1.0 / res["median_s"], "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": ("full Fock builds of the same workload (FP64 C/OpenMP oracle): 1 warm-up, "
                       "median of %d" % repeats),
            "seconds_median": res["median_s"], "seconds_all": res["times_s"],
            "nproc": nproc, "physical_cores": phys, "cpu_model": model,
            "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}}, err


class Build:
    """One workload's engine, resident densities and outputs; step() = one
    formGD-equivalent contraction list with the outputs zeroed first."""

    def __init__(self, name, nocc=20):
        from chronusq_public_amd import cqx
        self.mol, self.gbs, kind = workload(name)
        self.basis = cqx.BasisSet(self.gbs, self.mol)
        self.tpi = cqx.DirectTPI(self.basis, 1e-12, device=0)
        self.gto = cqx.GTODirectTPIContraction(self.tpi)
        nb = self.basis.nBasis
        self.mats, self.types, self.cplx = densities(kind, nb, nocc)
        self.dX = [cqx.DeviceMatrix(self.tpi, nb, self.cplx).upload(m) for m in self.mats]
        self.dA = [cqx.DeviceMatrix(self.tpi, nb, self.cplx) for _ in self.mats]
        self.cont = [cqx.TwoBodyContraction(x, a, True, t)
                     for x, a, t in zip(self.dX, self.dA, self.types)]
        # the first (warm-up) build times the lane counts of every class
        self.tpi.set_option("autotune", 1)
        self.tpi.set_option("streams", LAUNCH_STREAMS)

    def step(self):
        for a in self.dA:
            a.zero()
        self.gto.twoBodyContract(self.cont)

    def timed(self, steps, warmup):
        for _ in range(warmup):
            self.step()
        eri_ms, flops, f32 = [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
            st = self.tpi.last_stats()
            eri_ms.append(st["ms_eri"])
            flops.append(st["flops_model"])
            f32.append(st["flops_model_f32"])
        dt = time.perf_counter() - t0
        return dt, float(np.mean(eri_ms)), float(np.mean(flops)), float(np.mean(f32))

    def outputs(self):
        return [a.download() for a in self.dA]


def profile_counters(workload):
    """Counter figures of this workload's ERI launches from the committed
    rocprofv3 --pmc runs (profiles/eri_<workload>_counters.json): HBM
    traffic per build from copy-calibrated FETCH_SIZE / WRITE_SIZE, the
    executed FP64 FLOPs per build from SQ_INSTS_VALU_FLOPS_FP64 (calibrated
    on a kernel of known FMA count) and the memory-side atomic requests.
    Returned only when the file was measured on the sources being run (its
    source digest equals this tree's), else None."""
    from chronusq_public_amd.build import source_digest
    path = os.path.join(ROOT, "profiles", "eri_%s_counters.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d if d.get("source_digest") == source_digest() else None


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def mixed_comparison(steps, warmup, cut=1e-4):
    """FP64 and mixed builds of the alkane workload, timed alike; the mixed
    result is compared with the FP64 one (the FP64 path itself is held to
    1e-10 of the oracle by the GPU parity tests)."""
    progress("mixed precision: C40H82 / 6-31G* FP64 builds")
    b = Build("alkane40", nocc=161)
    dt64, eri64, fl64, _ = b.timed(steps, warmup)
    progress("mixed precision: mixed builds")
    J, K = b.outputs()
    st64 = b.tpi.last_stats()
    b.tpi.set_option("fp32_cut", cut)
    # the mixed warm-up build tunes the FP32 lists (and the shorter FP64 ones)
    b.tpi.set_option("autotune", 1)
    dtm, erim, flm, flm32 = b.timed(steps, warmup)
    Jm, Km = b.outputs()
    stm = b.tpi.last_stats()
    dG = (2 * Jm - Km) - (2 * J - K)
    D = b.mats[0]
    # roofline of the mixed build: FP64 part at the FP64 peak, FP32 part at the FP32 peak
    t_roof = (flm - flm32) / (FP64_VECTOR_PEAK * 1e12) + flm32 / (FP32_VECTOR_PEAK * 1e12)
    return {
        "workload": ("RHF J/K, C40H82 all-trans seed 0, 6-31G* (NB %d), D = C C^T with C NB x 161, "
                     "Schwarz 1e-12, FP32 when Q_ab Q_cd < %g" % (b.basis.nBasis, cut)),
        "steps": steps,
        "fp64_builds_per_s": steps / dt64,
        "mixed_builds_per_s": steps / dtm,
        "speedup": dt64 / dtm,
        "fp64_eri_ms": eri64,
        "mixed_eri_ms": erim,
        "quartets": stm["n_quartets"],
        "quartets_fp32": stm["n_quartets_f32"],
        "max_abs_dF": float(np.abs(dG).max()),
        "abs_dE": float(abs(0.25 * np.sum(D * dG))),
        "fp64_roofline_frac": fl64 / (eri64 * 1e-3) / (FP64_VECTOR_PEAK * 1e12),
        "mixed_roofline_frac": t_roof / (erim * 1e-3),
        "fp64_quartets_check": st64["n_quartets"] == stm["n_quartets"],
    }


def serving(njobs, max_inflight=8, cut=1e-4, seed=0):
    """SCF serving on water clusters (H2O)_n, n ~ U{4..48} from seed, in a
    cube of side 12 (n/16)^(1/3) A (the density of the GHF workload),
    cc-pVDZ, SAD guess, CDIIS(8), converged at RMS dD < 1e-8.  A closed-loop
    FP64 run (every job at t = 0) gives the peak throughput and the FP64
    energies; the measured run has Poisson arrivals at 80% of that rate and
    mixed-precision Fock builds (FP32 when Q_ab Q_cd < cut)."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = [int(n) for n in rng.integers(4, 49, size=njobs)]
    mols = [water_cluster(n, 12.0 * (n / 16.0) ** (1.0 / 3.0), seed=1000 + i) for i, n in enumerate(sizes)]
    progress(f"serving: closed-loop FP64 run of {njobs} jobs, sizes {sizes}")
    ref = cqx.serve([("cc-pVDZ", m, 0.0) for m in mols], max_inflight=max_inflight, verbose=1)
    peak = njobs / max(r["t_done"] for r in ref)
    lam = 0.8 * peak
    arrivals = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / lam, size=njobs - 1))])
    progress(f"serving: Poisson run at {lam:.3f} jobs/s, mixed precision")
    res = cqx.serve([("cc-pVDZ", m, float(a)) for m, a in zip(mols, arrivals)],
                    max_inflight=max_inflight, fp32_cut=cut, verbose=1)
    lat = np.array([r["t_done"] - r["t_arrival"] for r in res])
    dE = np.array([abs(r["e_total"] - q["e_total"]) for r, q in zip(res, ref)])
    progress("serving: the same Poisson run with cross-request batching of the Fock builds")
    resb = cqx.serve([("cc-pVDZ", m, float(a)) for m, a in zip(mols, arrivals)],
                     max_inflight=max_inflight, fp32_cut=cut, verbose=1, batch_fock=1)
    latb = np.array([r["t_done"] - r["t_arrival"] for r in resb])
    dEb = np.array([abs(r["e_total"] - q["e_total"]) for r, q in zip(resb, ref)])
    return {
        "workload": ("RHF SCF jobs, (H2O)_n n ~ U{4..48} (seed %d), cc-pVDZ (NB 96-1152), SAD guess, "
                     "CDIIS(8), RMS dD < 1e-8; Poisson arrivals at 0.8 x closed-loop peak; "
                     "mixed precision FP32 when Q_ab Q_cd < %g; %d slots" % (seed, cut, max_inflight)),
        "jobs": njobs,
        "sizes": sizes,
        "peak_jobs_per_s_fp64_closed_loop": peak,
        "arrival_rate_per_s": lam,
        "jobs_per_s": njobs / max(r["t_done"] for r in res),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "max_abs_dE_vs_fp64": float(dE.max()),
        "iterations_mean": float(np.mean([r["iterations"] for r in res])),
        "all_converged": all(r["converged"] for r in res) and all(r["converged"] for r in ref),
        "cross_request_batching": {
            "note": ("the same arrivals and jobs; every round builds the Fock matrices of all jobs in "
                     "flight in one batched contraction"),
            "jobs_per_s": njobs / max(r["t_done"] for r in resb),
            "p50_latency_s": float(np.percentile(latb, 50)),
            "p99_latency_s": float(np.percentile(latb, 99)),
            "max_abs_dE_vs_fp64": float(dEb.max()),
            "iterations_mean": float(np.mean([r["iterations"] for r in resb])),
            "all_converged": all(r["converged"] for r in resb),
        },
    }


def batched_comparison(nmol, steps, cut=1e-4, seed=0):
    """Batched uneven-size Fock builds: nmol water clusters (H2O)_n, n ~
    U{2..24} from seed, in cubes of side 12 (n/16)^(1/3) A, cc-pVDZ,
    D = C C^T with C NB x 5n, N(0, 1/sqrt NB).  The plain FP64 GPU build is
    one context per molecule, contracted one after the other; the batched
    build is one context over all molecules (every class pair one launch for
    the batch) with mixed precision (FP32 when Q_ab Q_cd < cut) and lane
    counts autotuned per class in the warm-up round.  A "build" is one Fock
    build of one molecule."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import water_cluster
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = [int(n) for n in rng.integers(2, 25, size=nmol)]
    progress(f"batched: {nmol} clusters, sizes {min(sizes)}..{max(sizes)}")
    bases, dens = [], []
    for i, n in enumerate(sizes):
        b = cqx.BasisSet("cc-pVDZ", water_cluster(n, 12.0 * (n / 16.0) ** (1.0 / 3.0), seed=2000 + i))
        C = rng.normal(0, 1 / np.sqrt(b.nBasis), size=(b.nBasis, 5 * n))
        D = C @ C.T
        bases.append(b)
        dens.append(0.5 * (D + D.T))

    def make(tpi, nb, D):
        X = cqx.DeviceMatrix(tpi, nb).upload(D)
        J, K = cqx.DeviceMatrix(tpi, nb), cqx.DeviceMatrix(tpi, nb)
        return X, J, K

    # plain FP64: one context per molecule, sequential
    singles = []
    for b, D in zip(bases, dens):
        tpi = cqx.DirectTPI(b, 1e-12, device=0)
        tpi.set_option("autotune", 1)  # tuned in the warm-up round, as the batched context
        X, J, K = make(tpi, b.nBasis, D)
        singles.append((tpi, cqx.GTODirectTPIContraction(tpi), X, J, K))

    def round_single():
        for tpi, gto, X, J, K in singles:
            J.zero()
            K.zero()
            gto.twoBodyContract([cqx.TwoBodyContraction(X, J, True, COULOMB_),
                                 cqx.TwoBodyContraction(X, K, True, EXCHANGE_)])

    round_single()
    t0 = time.perf_counter()
    for _ in range(steps):
        round_single()
    t_single = (time.perf_counter() - t0) / steps
    ref = [(J.download(), K.download()) for _, _, _, J, K in singles]
    # batched, mixed precision
    progress("batched: one context over all molecules, mixed precision")
    bb = cqx.BasisSet.batch(bases)
    tpi = cqx.DirectTPI(bb, 1e-12, device=0)
    tpi.set_option("fp32_cut", cut)
    tpi.set_option("autotune", 1)  # the warm-up round picks the lanes per class
    gto = cqx.GTODirectTPIContraction(tpi)
    mats = [make(tpi, b.nBasis, D) for b, D in zip(bases, dens)]
    cont = []
    for X, J, K in mats:
        cont += [cqx.TwoBodyContraction(X, J, True, COULOMB_), cqx.TwoBodyContraction(X, K, True, EXCHANGE_)]

    def round_batch():
        for _, J, K in mats:
            J.zero()
            K.zero()
        gto.twoBodyContract(cont)

    round_batch()
    t0 = time.perf_counter()
    for _ in range(steps):
        round_batch()
    t_batch = (time.perf_counter() - t0) / steps
    st = tpi.last_stats()
    dF, dE = 0.0, 0.0
    for (_, J, K), (J0, K0), D in zip(mats, ref, dens):
        dG = (2 * J.download() - K.download()) - (2 * J0 - K0)
        dF = max(dF, float(np.abs(dG).max()))
        dE = max(dE, float(abs(0.25 * np.sum(D * dG))))
    del gto, tpi, mats, cont
    # the same batched mixed build recorded once into a HIP graph (warm-up
    # build) and replayed: one list batch, so the lists of every molecule
    # stay resident in HBM (8 B per candidate and precision)
    progress("batched: recorded build (HIP graph) replayed, mixed precision")
    tg = cqx.DirectTPI(bb, 1e-12, device=0)
    tg.set_option("fp32_cut", cut)
    tg.set_option("list_cap", 2.0 ** 33)
    tg.set_option("autotune", 1)
    tg.set_option("graph", 1)
    gg = cqx.GTODirectTPIContraction(tg)
    gmats = [make(tg, b.nBasis, D) for b, D in zip(bases, dens)]
    gcont = []
    for X, J, K in gmats:
        gcont += [cqx.TwoBodyContraction(X, J, True, COULOMB_), cqx.TwoBodyContraction(X, K, True, EXCHANGE_)]

    def round_graph():
        for _, J, K in gmats:
            J.zero()
            K.zero()
        gg.twoBodyContract(gcont)

    round_graph()
    recorded = tg.last_stats()["graph"]
    t0 = time.perf_counter()
    for _ in range(steps):
        round_graph()
    t_graph = (time.perf_counter() - t0) / steps
    sg = tg.last_stats()
    gF, gE = 0.0, 0.0
    for (_, J, K), (J0, K0), D in zip(gmats, ref, dens):
        dG = (2 * J.download() - K.download()) - (2 * J0 - K0)
        gF = max(gF, float(np.abs(dG).max()))
        gE = max(gE, float(abs(0.25 * np.sum(D * dG))))
    return {
        "workload": ("RHF J/K of %d water clusters (H2O)_n, n ~ U{2..24} (seed %d), cc-pVDZ, "
                     "D = C C^T with C NB x 5n; batched = one context and one launch per class pair for "
                     "all molecules, FP32 when Q_ab Q_cd < %g" % (nmol, seed, cut)),
        "molecules": nmol,
        "fp64_single_builds_per_s": nmol / t_single,
        "batched_mixed_builds_per_s": nmol / t_batch,
        "speedup": t_single / t_batch,
        "batched_ms_eri": st["ms_eri"],
        "batched_ms_total": st["ms_total"],
        "quartets": st["n_quartets"],
        "quartets_fp32": st["n_quartets_f32"],
        "max_abs_dF": dF,
        "max_abs_dE": dE,
        "graph": {
            "batched_mixed_builds_per_s": nmol / t_graph,
            "speedup_vs_fp64_single": t_single / t_graph,
            "recorded": recorded == 1,
            "replayed": sg["graph"] == 2,
            "ms_graph": sg["ms_eri"],
            "ms_total": sg["ms_total"],
            "max_abs_dF": gF,
            "max_abs_dE": gE,
        },
    }


def host_leg(bench, steps, out, line):
    """The drop-in host-pointer path: X and AX in host memory (mem =
    CQX_HOST), as the ChronusQ adapter calls the engine; PCIe copies
    included, so never `value`."""
    from chronusq_public_amd import cqx
    progress(f"host-pointer builds (1 warmup + {steps} timed)")
    mats, types, cplx = bench.mats, bench.types, bench.cplx
    hmats = [np.asfortranarray(m, dtype=np.complex128 if cplx else np.float64) for m in mats]
    hout = [np.zeros_like(m, order="F") for m in hmats]
    hcont = [cqx.TwoBodyContraction(x, a, True, t) for x, a, t in zip(hmats, hout, types)]
    for a in hout:
        a[...] = 0
    bench.gto.twoBodyContract(hcont)
    t0 = time.perf_counter()
    for _ in range(steps):
        for a in hout:
            a[...] = 0
        bench.gto.twoBodyContract(hcont)
    hdt = time.perf_counter() - t0
    line["host_pointer"] = {
        "builds_per_s": steps / hdt,
        "ms_per_step": 1e3 * hdt / steps,
        "bytes_per_build_over_pcie": int(2 * sum(m.nbytes for m in hmats)),
        "max_abs_d_vs_device": float(max(np.abs(a - b).max() for a, b in zip(hout, out))),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ghf16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-mols", type=int, default=256,
                    help="molecules of the batched-build measurement (0 skips it)")
    ap.add_argument("--serve-jobs", type=int, default=24,
                    help="SCF jobs of the serving measurement (0 skips it)")
    ap.add_argument("--mixed-steps", type=int, default=3,
                    help="timed builds per precision of the mixed-precision comparison (0 skips it)")
    ap.add_argument("--opt", action="append", default=[], metavar="NAME=VALUE",
                    help="engine option of the main workload's context (cqx_set_option), repeatable")
    ap.add_argument("--streams", type=int, default=4,
                    help="streams the ERI launches of a build are dealt over, 1-4 (Fock-build "
                         "contexts of the main and mixed-precision legs)")
    ap.add_argument("--no-graph", action="store_true", help="skip the recorded-build (HIP graph) leg")
    ap.add_argument("--no-host", action="store_true", help="skip the host-pointer leg")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM "
                         "counter calibration for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        raise SystemExit("bench.py measures one GPU: the direct J/K engine runs one process "
                         "on one MI355X (multi-GPU sharding is not part of this build)")

    from chronusq_public_amd import cqx
    global LAUNCH_STREAMS
    LAUNCH_STREAMS = args.streams
    bench = Build(args.workload, nocc=161 if args.workload == "alkane40" else 20)
    for o in args.opt:
        name, value = o.split("=", 1)
        bench.tpi.set_option(name, float(value))
    mol, gbs, tpi = bench.mol, bench.gbs, bench.tpi
    mats, types, cplx = bench.mats, bench.types, bench.cplx
    nb = bench.basis.nBasis

    if args.copy_mb:
        import ctypes as C
        nbytes = args.copy_mb << 20
        src, dst = C.c_void_p(), C.c_void_p()
        cqx._check(cqx.lib().cqx_device_alloc(tpi._h, nbytes, C.byref(src)), tpi._h)
        cqx._check(cqx.lib().cqx_device_alloc(tpi._h, nbytes, C.byref(dst)), tpi._h)
        cqx._check(cqx.lib().cqx_memcpy(tpi._h, dst, src, nbytes, 2), tpi._h)
        cqx.lib().cqx_device_free(tpi._h, src)
        cqx.lib().cqx_device_free(tpi._h, dst)
    progress(f"{args.workload}: {args.warmup} warmup + {args.steps} timed builds")
    dt, eri_avg, flops_avg, _ = bench.timed(args.steps, args.warmup)
    st = tpi.last_stats()
    ms_step = 1e3 * dt / args.steps
    achieved = flops_avg / (eri_avg * 1e-3) / 1e12
    out = bench.outputs()

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
            "bound": "fp64 valu",
            "compute": "fp64 vector ALU (MI355X fp64 vector and matrix peaks are both 78.6 TFLOP/s)",
            "kernel": "ERI + J/K digestion launches of one build (all angular-momentum classes)",
            "achieved": achieved,
            "peak": FP64_VECTOR_PEAK,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK,
            "traffic": None,
            "flops_per_build": flops_avg,
            "eri_ms_per_build": eri_avg,
            # cross-check of the flop model: primitive quartets of the evaluated
            # shell quartets (Ka Kb Kc Kd, before primitive-pair screening) per second
            "prim_quartets_per_build": st["n_prim_quartets_max"],
            "prim_quartets_per_s": st["n_prim_quartets_max"] / (eri_avg * 1e-3),
        },
    }
    pc = profile_counters(args.workload)
    if pc:
        r = line["roofline"]
        r["traffic"] = pc["traffic_bytes_per_build"]
        r["traffic_note"] = ("HBM bytes per build of the ERI/digestion launches, copy-calibrated "
                             "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (%s)" % pc["summary"])
        r["executed_flops_per_build"] = pc["executed_fp64_flops_per_build"]
        r["executed_frac"] = pc["executed_fp64_flops_per_build"] / (eri_avg * 1e-3) / (FP64_VECTOR_PEAK * 1e12)
        r["memory_side_atomics_per_build"] = pc["tcc_ea0_atomic_per_build"]
        r["limiter"] = pc["limiter"]
        r["profile_source_digest"] = pc["source_digest"]
    else:
        line["roofline"]["traffic_note"] = ("no rocprofv3 counter profile of these sources: "
                                            "traffic and executed FLOPs omitted")
    # the drop-in host-pointer path (X and AX in host memory, mem = CQX_HOST,
    # as the ChronusQ adapter calls it): PCIe copies included, not `value`
    if not args.no_host:
        host_leg(bench, args.steps, out, line)
    # the same builds recorded once into a HIP graph and replayed (option
    # graph): the first build records, the timed ones replay; the outputs
    # are checked against the eager ones
    if not args.no_graph and "graph" not in [o.split("=", 1)[0] for o in args.opt]:
        progress(f"{args.workload}: recorded build, 1 warmup + {args.steps} timed replays")
        tpi.set_option("graph", 1)
        gdt, gms, _, _ = bench.timed(args.steps, 1)
        gst = tpi.last_stats()
        gout = bench.outputs()
        line["graph"] = {
            "builds_per_s": args.steps / gdt,
            "ms_per_step": 1e3 * gdt / args.steps,
            "ms_graph_per_build": gms,
            "replayed": gst["graph"] == 2,
            "max_abs_d_vs_eager": float(max(np.abs(a - b).max() for a, b in zip(gout, out))),
        }
        tpi.set_option("graph", 0)
    if args.mixed_steps > 0:
        line["mixed"] = mixed_comparison(args.mixed_steps, 1)
    if args.batch_mols > 0:
        line["batched"] = batched_comparison(args.batch_mols, 2)
    if args.serve_jobs > 0:
        line["serving"] = serving(args.serve_jobs)
    if not args.no_cpu:
        progress("cpu baseline: oracle builds on the host cores")
        cb, err = cpu_baseline(gbs, mol, mats, types, cplx, out)
        line["cpu_baseline"] = cb
        line["max_abs_dJK_vs_oracle"] = err
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
