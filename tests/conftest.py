import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_fixture(name):
    """(metadata dict, arrays) of a committed golden fixture."""
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    return meta, dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def fixture_molecule(meta):
    from chronusq_public_amd.chem import Molecule, basis_text
    mol = Molecule.from_angstrom(meta["geom_angstrom"])
    gbs = meta["basis_def"] or basis_text(meta["basis"])
    return mol, gbs


SCF_FIXTURES = ["water_631gd_rhf", "water_ccpvdz_unc_rhf_readden",
                "water_ccpvdz_unc_x2chf_readden", "b_sto3g_ghf",
                "kcakr_sto3g_ghf"]

# tests/func/contract.cxx:195-282 -- (contraction type, HER flag) used to
# produce each dataset, quirks included (NONHER/REAL/K is a COULOMB
# contraction, NONHER/COMPLEX/K a Hermitian EXCHANGE one).
CONTRACT_CASES = {
    "HER_REAL_J": (0, True), "HER_REAL_K": (1, True),
    "HER_COMPLEX_J": (0, True), "HER_COMPLEX_K": (1, True),
    "NONHER_REAL_J": (0, False), "NONHER_REAL_K": (0, False),
    "NONHER_COMPLEX_J": (0, False), "NONHER_COMPLEX_K": (1, True),
}
