"""Tile lists (option "tiles"): the lane-group class pairs of Hermitian
builds evaluate their quartets in tiles that share the bra's and the ket's
first shells, summing the shared J / K targets per tile.  Each build with
tiles = 1 on every lane-group class must equal the quartet-list build
(tiles = 0) of the same densities to 1e-12 relative, with the same quartets
(and FP32 quartets), for real RHF, complex two-component, mixed-precision,
batched and unscreened builds; the tile build also equals the oracle to
1e-10.  tiles = 2 takes tiles for a real J / K pair and quartet lists for
the eight exchange planes of a two-component build."""
import sys

import numpy as np
import pytest

from oracle.oracle import OracleBasis

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _density(rng, nb, nocc, cplx=False):
    C_ = rng.normal(0, 1 / np.sqrt(nb), size=(nb, nocc))
    if cplx:
        C_ = C_ + 1j * rng.normal(0, 1 / np.sqrt(nb), size=(nb, nocc))
    return np.asfortranarray(C_ @ C_.conj().T)


def _build(basis, Xs, types, tiles, screen=True, **opts):
    from chronusq_public_amd import cqx
    tpi = cqx.DirectTPI(basis, 1e-12)
    if tiles is not None:
        tpi.set_option("tiles", tiles)
    if tiles == 1:
        tpi.set_option("tile_classes", 255)
    for k, v in opts.items():
        tpi.set_option(k, v)
    gto = cqx.GTODirectTPIContraction(tpi)
    cplx = any(np.iscomplexobj(x) for x in Xs)
    dt = np.complex128 if cplx else np.float64
    AX = [np.zeros(x.shape, dtype=dt, order="F") for x in Xs]
    gto.twoBodyContract([cqx.TwoBodyContraction(np.asarray(x, dtype=dt), a, True, t)
                         for x, a, t in zip(Xs, AX, types)], screen=screen)
    return AX, tpi.last_stats()


def _same(A, B, sa, sb, tol=1e-12):
    for a, b in zip(A, B):
        scale = max(1.0, np.abs(b).max())
        assert np.abs(a - b).max() <= tol * scale, np.abs(a - b).max()
    assert sa["n_quartets"] == sb["n_quartets"] > 0
    assert sa["n_quartets_f32"] == sb["n_quartets_f32"]
    assert abs(sa["flops_model"] - sb["flops_model"]) <= 1e-9 * sb["flops_model"]


@pytest.mark.parametrize("case", ["rhf", "ghf", "unscreened"])
def test_tiles_equal_quartet_lists(case):
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import basis_text, water_cluster
    nw = 3 if case == "ghf" else 4
    mol = water_cluster(nw, 8.0, seed=7)
    gbs = basis_text("cc-pVDZ")
    b = cqx.BasisSet(gbs, mol)
    nb = b.nBasis
    rng = np.random.Generator(np.random.PCG64(11))
    say(f"case: tiles vs quartet lists, {case}, (H2O){nw}")
    if case == "ghf":
        D = _density(rng, nb, 5 * nw, cplx=True)
        Xs, types = [D, D, 0.3 * D, 0.2 * D.conj(), 0.1 * D], [0, 1, 1, 1, 1]
    else:
        D = _density(rng, nb, 5 * nw)
        Xs, types = [D, D], [0, 1]
    screen = case != "unscreened"
    A1, s1 = _build(b, Xs, types, 1, screen=screen)
    A0, s0 = _build(b, Xs, types, 0, screen=screen)
    _same(A1, A0, s1, s0)
    if case == "rhf":
        Jr, Kr = OracleBasis(gbs, mol.Z, mol.xyz).contract([D, D], [0, 1], [True, True])
        assert np.abs(A1[0] - Jr).max() < 1e-10 and np.abs(A1[1] - Kr).max() < 1e-10


def test_tiles_mixed_precision_and_bounded_lists():
    """C8H18 / 6-31G* with FP32 quartets below 1e-4: the same quartet and
    FP32 counts, and the same J / K to 1e-10 relative (the two paths may
    split a quartet's primitive sums over different lane counts, which
    rounds the FP32 quartets differently); a tiny list_cap puts the
    remaining class pairs in many batches."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import alkane_chain, basis_text
    mol = alkane_chain(8, 0)
    b = cqx.BasisSet(basis_text("6-31G*"), mol)
    rng = np.random.Generator(np.random.PCG64(3))
    D = _density(rng, b.nBasis, 33)
    say("case: tiles vs quartet lists, C8H18 mixed")
    A1, s1 = _build(b, [D, D], [0, 1], 1, fp32_cut=1e-4, list_cap=5000)
    A0, s0 = _build(b, [D, D], [0, 1], 0, fp32_cut=1e-4)
    assert s1["n_quartets_f32"] > 0
    _same(A1, A0, s1, s0, tol=1e-10)


def test_tiles_batched_molecules():
    """A batched context (three clusters of different sizes and one STO-3G
    molecule): every molecule's J / K and the quartet total equal the
    quartet-list build."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import basis_text, water_cluster
    parts = [cqx.BasisSet(basis_text("cc-pVDZ"), water_cluster(n, 9.0, seed=20 + n)) for n in (2, 4, 3)]
    parts.append(cqx.BasisSet(basis_text("STO-3G"), water_cluster(2, 6.0, seed=5)))
    bb = cqx.BasisSet.batch(parts)
    rng = np.random.Generator(np.random.PCG64(21))
    Xs, types = [], []
    for p in parts:
        D = _density(rng, p.nBasis, 6)
        Xs += [D, D]
        types += [0, 1]
    say("case: tiles vs quartet lists, batched")
    A1, s1 = _build(bb, Xs, types, 1)
    A0, s0 = _build(bb, Xs, types, 0)
    _same(A1, A0, s1, s0)


@pytest.mark.parametrize("case", ["rhf", "ghf"])
def test_tiles_2_takes_tiles_for_a_real_pair_only(case):
    """tiles = 2: an RHF J / K build goes through tile lists (its candidate
    count is the tile lists', not the quartet lists') and equals the
    quartet-list build; a GHF build with eight exchange planes stays on
    quartet lists (the same candidates), as does the default."""
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import basis_text, water_cluster
    b = cqx.BasisSet(basis_text("cc-pVDZ"), water_cluster(4, 8.0, seed=9))
    rng = np.random.Generator(np.random.PCG64(31))
    say(f"case: default tiles, {case}")
    if case == "ghf":
        D = _density(rng, b.nBasis, 20, cplx=True)
        Xs, types = [D, D, 0.3 * D, 0.2 * D.conj(), 0.1 * D], [0, 1, 1, 1, 1]
    else:
        D = _density(rng, b.nBasis, 20)
        Xs, types = [D, D], [0, 1]
    A2, s2 = _build(b, Xs, types, 2)
    A0, s0 = _build(b, Xs, types, 0)
    Ad, sd = _build(b, Xs, types, None)
    _same(A2, A0, s2, s0)
    assert sd["n_candidates"] == s0["n_candidates"]
    if case == "rhf":
        assert s2["n_candidates"] != s0["n_candidates"]
    else:
        assert s2["n_candidates"] == s0["n_candidates"]
