"""Nuclear-gradient J/K contraction on the GPU (DirectGradContraction::
directScaffoldGrad) against the oracle's restatement -- itself pinned by
finite differences of the forward contraction -- to 1e-10 per element for
every coordinate, the two-electron gradient of FockBuilder::getGDGrad, and
translational invariance; the unsupported cases fail with CErr."""
import sys

import numpy as np
import pytest

from conftest import fixture_molecule, load_fixture
from oracle.oracle import COULOMB, EXCHANGE, OracleBasis

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _cases():
    from chronusq_public_amd.chem import basis_text, water_cluster
    meta, d = load_fixture("scf_water_631gd_rhf")
    mol, gbs = fixture_molecule(meta)
    yield "water 6-31G(d) converged D", mol, gbs, d["PDM_SCALAR"], False
    yield "water 6-31G(d) Cartesian d", mol, gbs, None, True
    yield "(H2O)2 cc-pVDZ", water_cluster(2, 4.0, seed=5), basis_text("cc-pVDZ"), None, False
    # f shells: (ff|ff) and the other f classes, pure and Cartesian
    yield "water cc-pVTZ", water_cluster(1, 3.0, seed=1), basis_text("cc-pVTZ"), None, False
    yield "water cc-pVTZ Cartesian", water_cluster(1, 3.0, seed=1), basis_text("cc-pVTZ"), None, True
    # g shells: (gg|gg) needs the 2D tables one root at a time and 28 passes
    yield "water cc-pVQZ", water_cluster(1, 3.0, seed=1), basis_text("cc-pVQZ"), None, False


@pytest.mark.parametrize("idx", [0, 1, 2, 3, 4, 5])
def test_grad_contraction_matches_oracle(idx):
    from chronusq_public_amd import cqx
    name, mol, gbs, D, cart = list(_cases())[idx]
    say("case: gradient contraction", name)
    b = cqx.BasisSet(gbs, mol, force_cart=cart)
    tpi = cqx.DirectTPI(b, 1e-12, device=0)
    nb = b.nBasis
    if D is None:
        rng = np.random.Generator(np.random.PCG64(7 + idx))
        C_ = rng.normal(0, 1 / np.sqrt(nb), size=(nb, 6))
        D = C_ @ C_.T
    natom = len(mol.Z)
    grad, J, K = cqx.getGDGrad(D, tpi, natom)
    assert tpi.last_stats()["n_quartets"] > 0
    ob = OracleBasis(gbs, np.asarray(mol.Z, dtype=np.int32), np.asarray(mol.xyz), force_cart=cart)
    ref = ob.grad_contract([D, D], [COULOMB, EXCHANGE])
    for g in range(3 * natom):
        assert np.abs(J[g] - ref[g][0]).max() < 1e-10, (g, "J")
        assert np.abs(K[g] - ref[g][1]).max() < 1e-10, (g, "K")
    gref = np.array([0.25 * np.sum(D * (2 * ref[g][0] - ref[g][1])) for g in range(3 * natom)])
    assert np.abs(grad - gref).max() < 1e-10
    for x in range(3):  # translational invariance of the two-electron gradient
        assert abs(grad[x::3].sum()) < 1e-9


def test_grad_contraction_rejects_unsupported():
    from chronusq_public_amd import cqx
    from chronusq_public_amd.chem import basis_text, water_cluster
    say("case: gradient contraction errors")
    mol = water_cluster(1, 3.0, seed=1)
    # a batched context: gradient contractions take one molecule
    parts = [cqx.BasisSet(basis_text("cc-pVDZ"), mol), cqx.BasisSet(basis_text("STO-3G"), mol)]
    bt = cqx.DirectTPI(cqx.BasisSet.batch(parts), 1e-12, device=0)
    with pytest.raises(cqx.CErr):
        cqx.getGDGrad(np.eye(bt.basis.nBasis), bt, 6)
    dz = cqx.DirectTPI(cqx.BasisSet(basis_text("cc-pVDZ"), mol), 1e-12, device=0)
    nb = dz.basis.nBasis
    D = np.eye(nb)
    A = [np.zeros((nb, nb), order="F") for _ in range(9)]
    with pytest.raises(cqx.CErr):  # non-Hermitian (direct.hpp: "Nonhermetian NYI!")
        cqx.DirectGradContraction(dz).gradTwoBodyContract(
            [[cqx.TwoBodyContraction(D, A[g], False, cqx.COULOMB)] for g in range(9)])
