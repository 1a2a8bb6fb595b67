"""g shells (l = 4) on the GPU: water cc-pVQZ (g on O, f on H) through the
templated wave kernels (classes up to 1080 Cartesian components) and the
generic kernel, which builds quartets too large for LDS in passes over the
first shell's components and the roots (Rys rules up to nine roots).
Against the oracle (its g integrals pinned in test_oracle_g.py): the shell
pairs and Schwarz bounds (1e-12 relative), sampled contracted quartets of
every class up to (gg|gg) (1e-12 relative), and RHF, complex two-component,
non-Hermitian and Cartesian-g J / K (1e-10)."""
import sys

import numpy as np
import pytest

from oracle.oracle import COULOMB, EXCHANGE, OracleBasis

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _setup():
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import WATER, Molecule, basis_text
    mol = Molecule.from_angstrom(WATER)
    gbs = basis_text("cc-pVQZ")
    b = cqx.BasisSet(gbs, mol)
    tpi = cqx.DirectTPI(b, 1e-12)
    return cqx, mol, gbs, b, tpi, OracleBasis(gbs, mol.Z, mol.xyz)


def test_qz_pairs_and_schwarz():
    cqx, mol, gbs, b, tpi, ob = _setup()
    say("case: water cc-pVQZ pairs and Schwarz")
    s1, s2 = tpi.shell_pairs()
    o1, o2 = ob.shell_pairs()
    assert np.array_equal(s1, o1) and np.array_equal(s2, o2)
    Q, Qo = tpi.schwarz(), ob.schwarz()
    assert np.abs(Q - Qo).max() <= 1e-12 * Qo.max()


def test_qz_eri_quartets_every_class_through_g():
    cqx, mol, gbs, b, tpi, ob = _setup()
    say("case: water cc-pVQZ quartets")
    sh = ob.shells()
    by_l = {}
    for s, x in enumerate(sh):
        by_l.setdefault(x["l"], []).append(s)
    assert max(by_l) == 4
    rng = np.random.Generator(np.random.PCG64(2))
    quartets = []
    for la in by_l:
        for lb in by_l:
            for lc in by_l:
                for ld in by_l:
                    if max(la, lb, lc, ld) >= 3:
                        quartets.append(tuple(int(rng.choice(by_l[l])) for l in (la, lb, lc, ld)))
    g = by_l[4][0]
    quartets.append((g, g, g, g))
    got = tpi.eri_quartets(quartets)
    for q, v in zip(quartets, got):
        ref = ob.eri_quartet(*q)
        assert v.shape == ref.shape
        assert np.abs(v - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), q


@pytest.mark.parametrize("case", ["rhf", "ghf", "nonhermitian", "cartesian"])
def test_qz_contraction_against_oracle(case):
    cqx, mol, gbs, b, tpi, ob = _setup()
    if case == "cartesian":  # 15 Cartesian g functions: (gg|gg) of 50625 components
        b = cqx.BasisSet(gbs, mol, force_cart=True)
        tpi = cqx.DirectTPI(b, 1e-12)
        ob = OracleBasis(gbs, mol.Z, mol.xyz, force_cart=True)
    gto = cqx.GTODirectTPIContraction(tpi)
    nb = b.nBasis
    rng = np.random.Generator(np.random.PCG64(5))
    say("case: water cc-pVQZ contraction", case)
    if case == "ghf":
        C = rng.normal(0, 1 / np.sqrt(nb), (nb, 5)) + 1j * rng.normal(0, 1 / np.sqrt(nb), (nb, 5))
        D = C @ C.conj().T
        Xs, types, hers = [D, D, 0.3 * D, 0.2 * D.conj(), 0.1 * D], [COULOMB] + [EXCHANGE] * 4, [True] * 5
    elif case in ("rhf", "cartesian"):
        C = rng.normal(0, 1 / np.sqrt(nb), (nb, 5))
        D = C @ C.T
        Xs, types, hers = [D, D], [COULOMB, EXCHANGE], [True, True]
    else:
        X = rng.normal(0, 1 / nb, (nb, nb))
        Xs, types, hers = [X, X], [COULOMB, EXCHANGE], [False, False]
    Xs = [np.asfortranarray(x) for x in Xs]
    cplx = case == "ghf"
    dt = np.complex128 if cplx else np.float64
    AX = [np.zeros((nb, nb), dtype=dt, order="F") for _ in Xs]
    gto.twoBodyContract([cqx.TwoBodyContraction(x, a, h, t) for x, a, t, h in zip(Xs, AX, types, hers)])
    ref = ob.contract(Xs, types, hers)
    for a, r in zip(AX, ref):
        err = np.abs(a - r).max()
        say(f"  max |d| {err:.2e} (max |AX| {np.abs(r).max():.2e})")
        assert err < 1e-10
    assert tpi.last_stats()["n_quartets"] > 0


def test_qz_wave_and_generic_kernels_agree():
    """The templated wave kernels of the f and g classes up to 1080
    Cartesian components against the generic kernel for every class beyond
    the lane-group ones: complex Hermitian J and K and a non-Hermitian K of
    water cc-pVQZ, 1e-12 relative to the largest element."""
    cqx, mol, gbs, b, tpi, ob = _setup()
    say("case: water cc-pVQZ wave vs generic")
    gto = cqx.GTODirectTPIContraction(tpi)
    nb = b.nBasis
    rng = np.random.Generator(np.random.PCG64(9))
    C = rng.normal(0, 1 / np.sqrt(nb), (nb, 5)) + 1j * rng.normal(0, 1 / np.sqrt(nb), (nb, 5))
    D = np.asfortranarray(C @ C.conj().T)
    Xs, types, hers = [D, D, D], [COULOMB, EXCHANGE, EXCHANGE], [True, True, False]

    def run():
        AX = [np.zeros((nb, nb), dtype=np.complex128, order="F") for _ in Xs]
        gto.twoBodyContract([cqx.TwoBodyContraction(x, a, h, t) for x, a, t, h in zip(Xs, AX, types, hers)])
        return AX
    wave = run()
    tpi.set_option("use_wave", 0)
    gen = run()
    for w, g in zip(wave, gen):
        scale = np.abs(g).max()
        say(f"  wave vs generic {np.abs(w - g).max() / scale:.2e} rel")
        assert np.abs(w - g).max() < 1e-12 * scale


def test_qz_batched_mixed_and_recorded():
    """g shells through the other paths: a batched context of water
    cc-pVQZ and (H2O)2 cc-pVDZ gives each molecule its oracle J / K
    (1e-10); mixed precision keeps the classes with f or g shells in FP64
    and stays within 1e-7 of FP64 (max |dF|, dF = 2 dJ - dK); a recorded
    build replays to the eager result (1e-11)."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import WATER, Molecule, basis_text, water_cluster
    say("case: cc-pVQZ batched, mixed, recorded")
    mols = [(Molecule.from_angstrom(WATER), basis_text("cc-pVQZ")), (water_cluster(2, 4.0, seed=3), basis_text("cc-pVDZ"))]
    parts = [cqx.BasisSet(g, m) for m, g in mols]
    rng = np.random.Generator(np.random.PCG64(8))
    dens = []
    for p in parts:
        C = rng.normal(0, 1 / np.sqrt(p.nBasis), (p.nBasis, 5))
        dens.append(np.asfortranarray(C @ C.T))

    def run(tpi, Ds):
        gto = cqx.GTODirectTPIContraction(tpi)
        out, cont = [], []
        for D in Ds:
            J, K = np.zeros(D.shape, order="F"), np.zeros(D.shape, order="F")
            out.append((J, K))
            cont += [cqx.TwoBodyContraction(D, J, True, COULOMB), cqx.TwoBodyContraction(D, K, True, EXCHANGE)]
        gto.twoBodyContract(cont)
        return out
    bb = cqx.BasisSet.batch(parts)
    got = run(cqx.DirectTPI(bb, 1e-12), dens)
    for (m, g), D, (J, K) in zip(mols, dens, got):
        Jr, Kr = OracleBasis(g, m.Z, m.xyz).contract([D, D], [COULOMB, EXCHANGE], [True, True])
        assert np.abs(J - Jr).max() < 1e-10 and np.abs(K - Kr).max() < 1e-10
    # mixed precision on the QZ molecule alone
    t64 = cqx.DirectTPI(parts[0], 1e-12)
    (J64, K64), = run(t64, dens[:1])
    t32 = cqx.DirectTPI(parts[0], 1e-12)
    t32.set_option("fp32_cut", 1e-4)
    (J32, K32), = run(t32, dens[:1])
    st = t32.last_stats()
    assert 0 < st["n_quartets_f32"] < st["n_quartets"] == t64.last_stats()["n_quartets"]
    dF = np.abs((2 * J32 - K32) - (2 * J64 - K64)).max()
    say(f"  mixed max|dF| {dF:.2e}, FP32 quartets {st['n_quartets_f32']} of {st['n_quartets']}")
    assert dF < 1e-7
    # recorded build and its replays
    tg = cqx.DirectTPI(parts[0], 1e-12)
    tg.set_option("graph", 1)
    for it in range(3):
        (Jg, Kg), = run(tg, dens[:1])
        assert tg.last_stats()["graph"] == (1 if it == 0 else 2)
        assert np.abs(Jg - J64).max() < 1e-11 and np.abs(Kg - K64).max() < 1e-11
