"""ChronusQ binary-file interop on the GPU path: a density read from a
ChronusQ checkpoint drives one Fock build (READDEN) whose G = F - H matches
the same file's Fock and core Hamiltonian (1e-9, the contract bar), and an
RHF run saved in ChronusQ's layout diffs against ChronusQ's own file for
the same molecule within the SCF bars (energy 1e-8 Eh, matrices 1e-5)."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, fixture_molecule, load_fixture

pytestmark = pytest.mark.gpu
BIN = os.path.join(GOLDEN, "chronusq_bin")


def say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.mark.parametrize("name,fixture", [("water_6-31Gd.bin.ref", "scf_water_631gd_rhf"),
                                          ("B_swap_GHF_sto-3g.bin.ref", "scf_b_sto3g_ghf")])
def test_readden_fock_build_from_chronusq_file(name, fixture):
    from chronusq_public_amd import binfile as bf, cqx
    say("case: READDEN from", name)
    meta, _ = load_fixture(fixture)
    mol, gbs = fixture_molecule(meta)
    path = os.path.join(BIN, name)
    D = bf.read_density(path)
    with bf.BinFile(path) as f:
        F = f.read_pauli("/SCF/FOCK")
        H = f.read_pauli("/INTS/CORE_HAMILTONIAN")
    b = cqx.BasisSet(gbs, mol)
    gto = cqx.GTODirectTPIContraction(cqx.DirectTPI(b, 1e-12, device=0))
    ss = cqx.SingleSlaterState(D, gto)
    cqx.FockBuilder.formGD(ss)
    for i, (Fc, Hc) in enumerate(zip(F, H)):
        assert np.abs(ss.twoeH[i] - (Fc - Hc)).max() < 1e-9, (name, i)


def test_rhf_checkpoint_diffs_against_chronusq(tmp_path):
    from chronusq_public_amd import binfile as bf, cqx
    say("case: RHF water 6-31G(d) saved in ChronusQ layout")
    meta, _ = load_fixture("scf_water_631gd_rhf")
    mol, gbs = fixture_molecule(meta)
    b = cqx.BasisSet(gbs, mol)
    scf = cqx.RHF(cqx.DirectTPI(b, 1e-12, device=0))
    assert scf.SCF()["converged"]
    out = str(tmp_path / "water_6-31Gd.bin")
    bf.save_scf(out, scf, b)
    with bf.BinFile(out) as mine, bf.BinFile(os.path.join(BIN, "water_6-31Gd.bin.ref")) as ref:
        assert mine.read("/SCF/FIELD_TYPE").tolist() == ref.read("/SCF/FIELD_TYPE").tolist() == [1]
        assert abs(mine.read("/SCF/TOTAL_ENERGY")[0] - ref.read("/SCF/TOTAL_ENERGY")[0]) < 1e-8
        for e in ("ONE_BODY_ENERGY", "MANY_BODY_ENERGY"):
            assert abs(mine.read("/SCF/" + e)[0] - ref.read("/SCF/" + e)[0]) < 1e-5
        for ds in ("/SCF/1PDM_SCALAR", "/SCF/FOCK_SCALAR"):
            assert mine.dims(ds) == ref.dims(ds)
            assert np.abs(mine.read(ds) - ref.read(ds)).max() < 1e-5, ds
        for ds in ("/INTS/OVERLAP", "/INTS/KINETIC", "/INTS/POTENTIAL", "/INTS/CORE_HAMILTONIAN_SCALAR"):
            R = ref.read(ds)
            assert np.abs(mine.read(ds) - R).max() <= 1e-13 * max(1.0, np.abs(R).max()), ds
