"""The batched (many uneven water clusters) and serving (SCF jobs with
Poisson arrivals) workloads checked against the FP64 oracle, not only
against the engine run another way: a fault shared by the batched and the
single-molecule paths, or by a job in the mix and the same job alone, fails
here.

Bars: FP64 J and K within 1e-10 of the oracle (max abs); mixed precision
max |dF| <= 1e-7 and |dE| <= 1e-6 Eh (dF = 2 dJ - dK, dE = 1/4 Tr(D dF));
served FP64 jobs' F - H within 1e-9 of the oracle's G[D_final] and their
energy within 1e-8 Eh of the oracle's energy of D_final; mixed jobs
|dE| <= 1e-6 Eh.  The oracle restates the reference scaffold
(direct.hpp:1412-2123) and FockBuilder's G = 2J - K (fockbuilder/impl.hpp:
223-237)."""
import sys

import numpy as np
import pytest

from oracle.oracle import OracleBasis

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _contract(cqx, gto, dens, is_batch):
    out = []
    cont = []
    for D in dens:
        J = np.zeros(D.shape, order="F")
        K = np.zeros(D.shape, order="F")
        out.append((J, K))
        cont += [cqx.TwoBodyContraction(D, J, True, cqx.COULOMB), cqx.TwoBodyContraction(D, K, True, cqx.EXCHANGE)]
    gto.twoBodyContract(cont)
    return out


@pytest.mark.timeout(900)
def test_batched_prefix_against_oracle():
    """The first 12 clusters of the 256-cluster batched workload (cc-pVDZ,
    n ~ U{2..24} from seed 0, D = C C^T with C NB x 5n) in one batched
    context: every molecule's J and K against the oracle, in FP64 and with
    mixed precision (FP32 when Q_ab Q_cd < 1e-4)."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import basis_text, batched_clusters
    nmol = 12
    sizes, mols, density = batched_clusters(256)
    gbs = basis_text("cc-pVDZ")
    bases = [cqx.BasisSet(gbs, m) for m in mols[:nmol]]
    dens = [density(b.nBasis, n) for b, n in zip(bases, sizes[:nmol])]
    bb = cqx.BasisSet.batch(bases)
    res = {}
    for cut in (0.0, 1e-4):
        tpi = cqx.DirectTPI(bb, 1e-12)
        if cut:
            tpi.set_option("fp32_cut", cut)
        res[cut] = _contract(cqx, cqx.GTODirectTPIContraction(tpi), [np.asfortranarray(D) for D in dens], True)
        st = tpi.last_stats()
        say(f"case: batched prefix of {nmol}, fp32_cut {cut}: {st['n_quartets']} quartets, "
            f"{st['n_quartets_f32']} in FP32")
        if cut:
            assert st["n_quartets_f32"] > st["n_quartets"] // 4
    for i, (mol, D) in enumerate(zip(mols[:nmol], dens)):
        say(f"  case: oracle, molecule {i} (H2O)_{sizes[i]}")
        Jr, Kr = OracleBasis(gbs, mol.Z, mol.xyz).contract([D, D], [0, 1], [True, True])
        J, K = res[0.0][i]
        assert np.abs(J - Jr).max() <= 1e-10 and np.abs(K - Kr).max() <= 1e-10, i
        Jm, Km = res[1e-4][i]
        dF = (2 * Jm - Km) - (2 * Jr - Kr)
        assert np.abs(dF).max() <= 1e-7, i
        assert abs(0.25 * np.sum(D * dF)) <= 1e-6, i


@pytest.mark.timeout(900)
def test_batched_full_size_mixed_equals_fp64_singles():
    """All 256 clusters: one batched mixed-precision build equals the 256
    single-molecule FP64 builds within max |dF| <= 1e-7 and |dE| <= 1e-6 Eh
    per molecule, and evaluates exactly their quartets."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import basis_text, batched_clusters
    sizes, mols, density = batched_clusters(256)
    gbs = basis_text("cc-pVDZ")
    bases = [cqx.BasisSet(gbs, m) for m in mols]
    dens = [np.asfortranarray(density(b.nBasis, n)) for b, n in zip(bases, sizes)]
    say("case: 256 clusters, FP64 one by one")
    ref, nq = [], 0
    for b, D in zip(bases, dens):
        tpi = cqx.DirectTPI(b, 1e-12)
        ref.append(_contract(cqx, cqx.GTODirectTPIContraction(tpi), [D], False)[0])
        nq += tpi.last_stats()["n_quartets"]
    say("case: 256 clusters, one batched mixed build")
    tpi = cqx.DirectTPI(cqx.BasisSet.batch(bases), 1e-12)
    tpi.set_option("fp32_cut", 1e-4)
    got = _contract(cqx, cqx.GTODirectTPIContraction(tpi), dens, True)
    assert tpi.last_stats()["n_quartets"] == nq
    worst_f, worst_e = 0.0, 0.0
    for (J, K), (J0, K0), D in zip(got, ref, dens):
        dF = (2 * J - K) - (2 * J0 - K0)
        worst_f = max(worst_f, float(np.abs(dF).max()))
        worst_e = max(worst_e, abs(0.25 * float(np.sum(D * dF))))
    say(f"  max |dF| {worst_f:.2e}, max |dE| {worst_e:.2e}")
    assert worst_f <= 1e-7 and worst_e <= 1e-6


@pytest.mark.timeout(900)
@pytest.mark.parametrize("cut", [0.0, 1e-4])
@pytest.mark.parametrize("batch_fock", [0, 1])
def test_served_jobs_against_oracle(cut, batch_fock):
    """A Poisson mix of the serving workload's generator (cc-pVDZ water
    clusters, job i from seed 1000 + i; sizes U{4..14} here so that the
    oracle checks every job within the test's time) served with up to 4 jobs
    in flight: for every job the oracle recomputes G[D_final]; FP64 jobs
    match F - H to 1e-9 and the energy to 1e-8 Eh, mixed jobs the energy
    to 1e-6 Eh."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import basis_text, poisson_arrivals, serving_jobs
    sizes, mols, rng = serving_jobs(6, seed=0, nmin=4, nmax=14)
    arrivals = poisson_arrivals(rng, len(mols), 4.0)
    gbs = basis_text("cc-pVDZ")
    say(f"case: served mix sizes {sizes}, fp32_cut {cut}, batch_fock {batch_fock}")
    res = cqx.serve([("cc-pVDZ", m, float(a)) for m, a in zip(mols, arrivals)], max_inflight=4,
                    fp32_cut=cut, matrices=True, batch_fock=batch_fock)
    for i, (mol, r) in enumerate(zip(mols, res)):
        assert r["status"] == 0 and r["converged"], i
        b = cqx.BasisSet(gbs, mol)
        H = 2.0 * (b.one_body(cqx.KINETIC) + b.one_body(cqx.POTENTIAL))
        D, F = r["D"], r["F"]
        say(f"  case: oracle, job {i} (H2O)_{sizes[i]}")
        J, K = OracleBasis(gbs, mol.Z, mol.xyz).contract([D, D], [0, 1], [True, True])
        G = 2 * J - K
        E = 0.5 * np.sum(D * H) + 0.25 * np.sum(D * G) + mol.nuclear_repulsion()
        if cut == 0.0:
            assert np.abs((F - H) - G).max() <= 1e-9, i
            assert abs(E - r["e_total"]) <= 1e-8, i
        else:
            assert abs(E - r["e_total"]) <= 1e-6, i
