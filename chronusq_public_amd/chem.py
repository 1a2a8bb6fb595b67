"""Molecules, units and basis-set files for the direct J/K engine.

Units follow ChronusQ: geometry in Angstrom is divided by AngPerBohr =
0.52917721092 (include/physcon.hpp:32, src/cxxapi/input/molopts.cxx:179-182).
Basis names map to G94 files the way ChronusQ's table does
(include/basisset/basisset_def.hpp:235-244).
"""
import os

import numpy as np

ANG_PER_BOHR = 0.52917721092

ELEMENTS = [
    "X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg",
    "Al", "Si", "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn",
    "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb",
    "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In",
    "Sn", "Sb", "Te", "I", "Xe"]

BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
BASIS_FILES = {
    "STO-3G": "sto3g.gbs",
    "6-31G": "6-31g.gbs",
    "6-31G(D)": "6-31gs.gbs",
    "6-31G*": "6-31gs.gbs",
    "CC-PVDZ": "cc-pvdz.gbs",
    "CC-PVTZ": "cc-pvtz.gbs",
    "CC-PVQZ": "cc-pvqz.gbs",
}


def atomic_number(sym):
    for z, e in enumerate(ELEMENTS):
        if z and e.lower() == sym.lower():
            return z
    raise KeyError(sym)


def basis_text(name):
    """G94 text of a named basis set."""
    key = name.strip().upper()
    if key not in BASIS_FILES:
        raise KeyError(f"unknown basis {name!r}")
    with open(os.path.join(BASIS_DIR, BASIS_FILES[key])) as f:
        return f.read()


class Molecule:
    """Atoms as atomic numbers plus Cartesian coordinates in bohr."""

    def __init__(self, Z, xyz_bohr):
        self.Z = np.asarray(Z, dtype=np.int32)
        self.xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)

    @classmethod
    def from_angstrom(cls, atoms):
        """atoms: iterable of (symbol, x, y, z) in Angstrom."""
        Z = [atomic_number(a[0]) for a in atoms]
        xyz = np.array([[a[1], a[2], a[3]] for a in atoms], dtype=np.float64)
        return cls(Z, xyz / ANG_PER_BOHR)

    @property
    def natom(self):
        return len(self.Z)

    def nuclear_repulsion(self):
        e = 0.0
        for i in range(self.natom):
            for j in range(i):
                e += self.Z[i] * self.Z[j] / np.linalg.norm(self.xyz[i] - self.xyz[j])
        return e

    def nelectron(self, charge=0):
        return int(self.Z.sum()) - charge


# Rigid water monomer of the ChronusQ test decks (Angstrom),
# tests/scf/serial/rhf/water_6-31Gd.inp.
WATER = [("O", 0.0, -0.07579184359, 0.0),
         ("H", 0.866811829, 0.6014357793, 0.0),
         ("H", -0.866811829, 0.6014357793, 0.0)]


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(n, box_angstrom, seed, min_oo=2.6):
    """(H2O)_n: rigid monomers with uniform random orientation (unit
    quaternion) and O uniform in a cube of side box_angstrom, rejection
    sampled so every O-O distance is >= min_oo Angstrom.  numpy PCG64."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mono = np.array([a[1:] for a in WATER])
    mono = mono - mono[0]
    atoms, opos = [], []
    tries = 0
    while len(opos) < n:
        tries += 1
        if tries > 100000:
            raise RuntimeError("cannot place waters")
        R = _random_rotation(rng)
        o = rng.uniform(0.0, box_angstrom, size=3)
        if any(np.linalg.norm(o - p) < min_oo for p in opos):
            continue
        opos.append(o)
        xyz = mono @ R.T + o
        atoms += [("O", *xyz[0]), ("H", *xyz[1]), ("H", *xyz[2])]
    return Molecule.from_angstrom(atoms)


def alkane_chain(ncarbon, seed, jitter=0.05, cc=1.53, ch=1.09):
    """All-trans linear alkane C_nH_{2n+2} with tetrahedral angles and
    seeded uniform +-jitter Angstrom displacements of every atom."""
    rng = np.random.Generator(np.random.PCG64(seed))
    half = np.deg2rad(109.4712206 / 2.0)
    dx, dy = cc * np.sin(half), cc * np.cos(half)
    carbons = np.array([[i * dx, (dy / 2) * (1 if i % 2 else -1), 0.0]
                        for i in range(ncarbon)])
    atoms = [("C", *c) for c in carbons]
    hs = []
    for i, c in enumerate(carbons):
        # both chain neighbours lie on the far side in y: the H-C-H bisector
        # points away from them (+y for the upper carbons), the two H leave
        # the chain plane symmetrically, H-C-H tetrahedral
        up = np.array([0.0, 1.0 if i % 2 else -1.0, 0.0])
        for s in (1.0, -1.0):
            v = up * np.cos(half) + np.array([0.0, 0.0, s]) * np.sin(half)
            hs.append(c + ch * v)
    # the terminal carbons take a third H where the next chain carbon would be
    for i, sgn in ((0, -1.0), (ncarbon - 1, 1.0)):
        c = carbons[i]
        yside = 1.0 if i % 2 else -1.0
        v = np.array([sgn * np.sin(half), -yside * np.cos(half), 0.0])
        hs.append(c + ch * v)
    atoms += [("H", *h) for h in hs]
    mol = Molecule.from_angstrom(atoms)
    mol.xyz = mol.xyz + rng.uniform(-jitter, jitter, size=mol.xyz.shape) / ANG_PER_BOHR
    return mol


def cluster_box(n):
    """Cube side (Angstrom) for (H2O)_n at the density of the 16-water,
    12 A cluster: 12 (n / 16)^(1/3)."""
    return 12.0 * (n / 16.0) ** (1.0 / 3.0)


def batched_clusters(nmol, seed=0, nmin=2, nmax=24):
    """The batched-build workload: nmol water clusters (H2O)_n, n ~
    U{nmin..nmax} from seed (numpy PCG64), cluster i placed from seed
    2000 + i in a cube of cluster_box(n); the densities D = C C^T,
    symmetrised, with C of NB x 5n entries N(0, 1/sqrt(NB)) drawn after the
    sizes from the same generator (NB = 24 n for cc-pVDZ).  Returns
    (sizes, molecules, a function nb -> density drawing the next one)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = [int(n) for n in rng.integers(nmin, nmax + 1, size=nmol)]
    mols = [water_cluster(n, cluster_box(n), seed=2000 + i) for i, n in enumerate(sizes)]

    def density(nb, n):
        C = rng.normal(0, 1 / np.sqrt(nb), size=(nb, 5 * n))
        D = C @ C.T
        return 0.5 * (D + D.T)

    return sizes, mols, density


def serving_jobs(njobs, seed=0, nmin=4, nmax=48):
    """The serving workload: njobs water clusters (H2O)_n, n ~ U{nmin..nmax}
    from seed (numpy PCG64), job i placed from seed 1000 + i in a cube of
    cluster_box(n).  Returns (sizes, molecules, rng) -- the generator then
    draws the Poisson arrival gaps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = [int(n) for n in rng.integers(nmin, nmax + 1, size=njobs)]
    mols = [water_cluster(n, cluster_box(n), seed=1000 + i) for i, n in enumerate(sizes)]
    return sizes, mols, rng


def poisson_arrivals(rng, njobs, rate):
    """Arrival times of njobs jobs: the first at 0, then exponential gaps
    of mean 1 / rate."""
    return np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / rate, size=njobs - 1))])
