#!/usr/bin/env python3
"""Extract ChronusQ's own golden J/K and Fock data into small numpy fixtures.

Runs only where the ChronusQ source tree is mounted (it reads HDF5 files
under ``/root/reference/tests``) and where ``h5dump`` is available.  The
outputs are committed; nothing at test time reads the reference tree.

Sources (all data written by ChronusQ's DIRECT J/K path):

* ``tests/func/reference/contract.hdf5`` -- the X -> AX pairs that
  ``tests/func/contract.cxx:97-182`` checks against
  ``GTODirectTPIContraction::twoBodyContract`` (water / cc-pVTZ, NB = 58).
* ``tests/scf/reference/*.bin.ref`` -- SCF restart files holding
  ``/SCF/FOCK_*``, ``/INTS/CORE_HAMILTONIAN_*`` and ``/SCF/1PDM_*``, so that
  G[D] = F - H is ChronusQ's own ``FockBuilder::formGD`` output
  (``include/fockbuilder/impl.hpp:75-244``); also ``/INTS/KINETIC``,
  ``/INTS/POTENTIAL`` and the Lowdin ``/SCF/ORTHO`` where the file has them.  The molecule and basis of each
  fixture are read from the input deck that produced it.

Matrices are stored as mathematical matrices ``M[p, q]``.  ChronusQ writes
its column-major buffers straight to HDF5, so the HDF5 (row-major) view is
the transpose; the transpose is undone here.
"""
import json
import os
import re
import shutil
import subprocess
import sys

import numpy as np

REF = "/root/reference"
H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))

BIN_FILES = ("water_6-31Gd.bin.ref", "B_swap_GHF_sto-3g.bin.ref")

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|[-+]?nan|[-+]?inf"


def h5_dataset(path, name):
    """Return a dataset as numpy (complex for r/i compounds), HDF5 shape."""
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", name, path],
                         check=True, capture_output=True, text=True).stdout
    is_complex = "H5T_COMPOUND" in out
    m = re.search(r"DATASPACE\s+(SIMPLE|SCALAR)\s*(?:\{\s*\(([^)]*)\))?", out)
    shape = () if m.group(1) == "SCALAR" else tuple(
        int(x) for x in m.group(2).split(","))
    body = out[out.index("DATA {") + len("DATA {"):]
    # drop the "(i,j):" index prefixes before collecting the numbers
    body = re.sub(r"\(\s*\d+(?:\s*,\s*\d+)*\s*\)\s*:", " ", body)
    vals = [float(v) for v in re.findall(NUM, body)]
    arr = np.array(vals, dtype=np.float64)
    if is_complex:
        arr = arr[0::2] + 1j * arr[1::2]
    return arr.reshape(shape)


def colmajor(a):
    """HDF5 view of a column-major ChronusQ matrix -> math matrix."""
    return np.ascontiguousarray(a.T)


def parse_deck(path):
    """Pull molecule, basis and reference type out of a ChronusQ input deck."""
    text = open(path).read()
    geom, basis, basis_def, ref = [], None, None, None
    lines = text.splitlines()
    i = 0
    section = None
    while i < len(lines):
        raw = lines[i]
        line = raw.split("#")[0].rstrip()
        s = line.strip()
        if s.startswith("[") and s.endswith("]"):
            section = s[1:-1].upper()
            i += 1
            continue
        low = s.lower().replace(" ", "")
        if section == "MOLECULE" and low.startswith("geom:"):
            i += 1
            while i < len(lines):
                t = lines[i].split()
                if len(t) != 4 or lines[i].strip().startswith("["):
                    break
                # isotope labels (Ne-22) name the element for the electrons
                geom.append([t[0].split("-")[0], float(t[1]), float(t[2]), float(t[3])])
                i += 1
            continue
        if section == "QM" and low.startswith("reference="):
            ref = s.split("=", 1)[1].strip()
        if section == "BASIS" and low.startswith("basis="):
            basis = s.split("=", 1)[1].strip()
        if section == "BASIS" and low.startswith("basisdef:"):
            i += 1
            block = []
            while i < len(lines) and not lines[i].strip().startswith("["):
                block.append(lines[i].strip())
                i += 1
            basis_def = "\n".join(block).strip() + "\n"
            continue
        i += 1
    return {"geom_angstrom": geom, "basis": basis, "basis_def": basis_def,
            "reference": ref}


CONTRACT = os.path.join(REF, "tests/func/reference/contract.hdf5")

# fixture name -> (bin.ref file, input deck)
SCF = {
    "water_ccpvdz_unc_rhf_readden": (
        "water_cc-pVDZ_RHF_RHFGuess_READDEN_scratch.bin.ref",
        "scf/serial/rhf/water_cc-pVDZ_RHF_READDEN.inp"),
    "water_ccpvdz_unc_x2chf_readden": (
        "water_cc-pVDZ_X2CHF_X2CHFGuess_READDEN_scratch.bin.ref",
        "scf/serial/x2c/water_cc-pVDZ_X2CHF_READDEN.inp"),
    "water_631gd_rhf": (
        "water_6-31Gd.bin.ref", "scf/serial/rhf/water_6-31Gd.inp"),
    "b_sto3g_ghf": (
        "B_swap_GHF_sto-3g.bin.ref", "scf/serial/ghf/B_swap_GHF_sto-3g.inp"),
    "kcakr_sto3g_ghf": (
        "KCaKr_sto-3G.bin.ref", "scf/serial/ghf/KCaKr_sto-3G.inp"),
    "nehe_sto3g_rhf": (
        "NeHe_rhf_sto-3g_iso.bin.ref", "scf/serial/rhf/NeHe_rhf_sto-3g_iso.inp"),
}

COMPS = ["SCALAR", "MZ", "MY", "MX"]


def main():
    if not os.path.exists(CONTRACT):
        sys.exit("ChronusQ source tree not mounted; fixtures are committed")

    out = {}
    for her in ("HER", "NONHER"):
        for fld in ("REAL", "COMPLEX"):
            for kind in ("J", "K"):
                for d in ("X", "AX"):
                    a = h5_dataset(CONTRACT, f"/CONTRACTION/{her}/{fld}/{kind}/{d}")
                    out[f"{her}_{fld}_{kind}_{d}"] = colmajor(a)
    np.savez_compressed(os.path.join(HERE, "contract_water_ccpvtz.npz"), **out)
    deck = parse_deck(os.path.join(REF, "tests/func/input/contract_ref.inp"))
    with open(os.path.join(HERE, "contract_water_ccpvtz.json"), "w") as f:
        json.dump({"source": "tests/func/reference/contract.hdf5",
                   "deck": "tests/func/input/contract_ref.inp",
                   "checked_by": "tests/func/contract.cxx:97-182 (tol 1e-9)",
                   **deck}, f, indent=1)

    for name, (ref, inp) in SCF.items():
        path = os.path.join(REF, "tests/scf/reference", ref)
        listing = subprocess.run([H5DUMP, "-n", path], check=True,
                                 capture_output=True, text=True).stdout
        have = set(re.findall(r"dataset\s+(\S+)", listing))
        arrs = {}
        for c in COMPS:
            for grp, key in (("/SCF/FOCK_", "FOCK_"), ("/SCF/1PDM_", "PDM_"),
                             ("/INTS/CORE_HAMILTONIAN_", "HCORE_")):
                ds = grp + c
                if ds in have:
                    arrs[key + c] = colmajor(h5_dataset(path, ds))
        arrs["OVERLAP"] = colmajor(h5_dataset(path, "/INTS/OVERLAP"))
        # one-electron pieces and the Lowdin transformation, where written
        for ds, key in (("/INTS/KINETIC", "KINETIC"), ("/INTS/POTENTIAL", "POTENTIAL"),
                        ("/SCF/ORTHO", "ORTHO")):
            if ds in have:
                arrs[key] = colmajor(h5_dataset(path, ds))
        for e in ("TOTAL_ENERGY", "MANY_BODY_ENERGY", "ONE_BODY_ENERGY"):
            arrs[e] = h5_dataset(path, "/SCF/" + e).reshape(-1)
        np.savez_compressed(os.path.join(HERE, f"scf_{name}.npz"), **arrs)
        deck = parse_deck(os.path.join(REF, "tests", inp))
        with open(os.path.join(HERE, f"scf_{name}.json"), "w") as f:
            json.dump({"source": "tests/scf/reference/" + ref,
                       "deck": "tests/" + inp, **deck}, f, indent=1)
        print(name, {k: v.shape for k, v in arrs.items()})

    # two ChronusQ binary files kept whole (data files of the reference's
    # own tests) for the binary-file interop: a real RHF file and a complex
    # GHF one, read there with this package's HDF5 reader and checked
    # against the h5dump extractions above
    os.makedirs(os.path.join(HERE, "chronusq_bin"), exist_ok=True)
    for ref in BIN_FILES:
        shutil.copyfile(os.path.join(REF, "tests/scf/reference", ref), os.path.join(HERE, "chronusq_bin", ref))


if __name__ == "__main__":
    main()
