"""Python mirror of ChronusQ's direct J/K interface over the cqx C ABI.

Names and argument meaning follow the reference classes:

* ``TwoBodyContraction``  -- include/particleintegrals/twopints.hpp:59-75
* ``BasisSet``            -- include/basisset/basisset_def.hpp:120-223
* ``DirectTPI``           -- include/particleintegrals/twopints/gtodirecttpi.hpp:30-141
* ``GTODirectTPIContraction.twoBodyContract``
                          -- gtodirecttpi.hpp:143-228 (-> direct.hpp:1412-2123)
* ``FockBuilder.formGD``  -- include/fockbuilder/impl.hpp:75-244

Errors raise ``CErr`` (a RuntimeError), as ChronusQ's CErr throws
std::runtime_error("FATAL") (include/cerr.hpp:33-51).  Every call goes to
libcqx.so on the GPU; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .chem import Molecule, basis_text

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libcqx.so")

COULOMB, EXCHANGE = 0, 1
HOST, DEVICE = 0, 1


class CErr(RuntimeError):
    """ChronusQ's fatal error."""


_lib = None


class _Contraction(C.Structure):
    _fields_ = [("X", C.c_void_p), ("AX", C.c_void_p), ("her", C.c_int), ("type", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("n_candidates", C.c_longlong), ("n_quartets", C.c_longlong),
                ("n_prim_quartets_max", C.c_longlong), ("flops_model", C.c_double),
                ("ms_total", C.c_double), ("ms_eri", C.c_double),
                ("n_launch_eri", C.c_int), ("n_quartets_f32", C.c_longlong),
                ("flops_model_f32", C.c_double), ("graph", C.c_int)]


SYMBOLS = ["cqx_basis_create", "cqx_basis_from_g94", "cqx_basis_destroy",
           "cqx_basis_dims", "cqx_basis_shell", "cqx_basis_shell_prims",
           "cqx_ctx_create", "cqx_ctx_destroy", "cqx_last_error", "cqx_ctx_schwarz",
           "cqx_ctx_shell_pairs", "cqx_two_body_contract", "cqx_eri_quartets",
           "cqx_last_stats", "cqx_set_option", "cqx_device_alloc", "cqx_device_free",
           "cqx_memcpy", "cqx_device_zero", "cqx_one_body", "cqx_basis_batch", "cqx_scf_default_opts",
           "cqx_scf_create", "cqx_scf_iterate", "cqx_scf_run", "cqx_scf_result_get",
           "cqx_scf_matrix", "cqx_scf_orbital_energies", "cqx_scf_destroy",
           "cqx_scf_last_error", "cqx_serve_default_opts", "cqx_serve", "cqx_grad_two_body_contract"]

OVERLAP, KINETIC, POTENTIAL = 0, 1, 2


def lib():
    """Load libcqx.so; fails loudly when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise CErr(f"{LIBPATH} is missing: build it with chronusq_public_amd/build.py")
        L = C.CDLL(LIBPATH)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.cqx_last_error.restype = C.c_char_p
        L.cqx_last_error.argtypes = [P]
        L.cqx_basis_from_g94.argtypes = [C.c_char_p, I, P, P, I, C.POINTER(P)]
        L.cqx_basis_create.argtypes = [I, P, P, P, P, P, P, P, C.POINTER(P)]
        L.cqx_basis_destroy.argtypes = [P]
        L.cqx_basis_dims.argtypes = [P, P, P, P, P]
        L.cqx_basis_shell.argtypes = [P, I, P, P, P, P, P]
        L.cqx_basis_shell_prims.argtypes = [P, I, P, P, P]
        L.cqx_ctx_create.argtypes = [P, D, I, C.POINTER(P)]
        L.cqx_ctx_destroy.argtypes = [P]
        L.cqx_ctx_schwarz.argtypes = [P, P]
        L.cqx_ctx_shell_pairs.argtypes = [P, P, P, P]
        L.cqx_two_body_contract.argtypes = [P, I, I, P, I, I]
        L.cqx_eri_quartets.argtypes = [P, I, P, P, P, P, P, P]
        L.cqx_last_stats.argtypes = [P, C.POINTER(Stats)]
        L.cqx_set_option.argtypes = [P, C.c_char_p, D]
        L.cqx_device_alloc.argtypes = [P, C.c_size_t, C.POINTER(P)]
        L.cqx_device_free.argtypes = [P, P]
        L.cqx_memcpy.argtypes = [P, P, P, C.c_size_t, I]
        L.cqx_device_zero.argtypes = [P, P, C.c_size_t]
        L.cqx_one_body.argtypes = [P, I, I, P, P, P]
        L.cqx_basis_batch.argtypes = [I, P, C.POINTER(P)]
        L.cqx_grad_two_body_contract.argtypes = [P, I, I, P, I]
        _lib = L
    return _lib


def _check(rc, handle=None):
    if rc != 0:
        msg = lib().cqx_last_error(handle)
        raise CErr(f"FATAL (cqx status {rc}): {msg.decode() if msg else ''}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class TwoBodyContraction:
    """{X, AX, HER, contType} (twopints.hpp:59-75).  X / AX are NB x NB
    numpy arrays (column-major semantics: element (i, j) is M[i, j]) or
    device tensors exposing ``data_ptr()``."""

    def __init__(self, X, AX, HER=True, contType=COULOMB):
        self.X, self.AX, self.HER, self.contType = X, AX, bool(HER), int(contType)


class BasisSet:
    """Shell set on a molecule (ReferenceBasisSet + BasisSet::update)."""

    def __init__(self, name_or_g94, mol, force_cart=False):
        text = name_or_g94 if "****" in name_or_g94 else basis_text(name_or_g94)
        self.mol = mol
        Z = np.ascontiguousarray(mol.Z, dtype=np.int32)
        xyz = np.ascontiguousarray(mol.xyz, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().cqx_basis_from_g94(text.encode(), len(Z), _ptr(Z), _ptr(xyz),
                                        int(force_cart), C.byref(h)))
        self._h = h
        dims = [C.c_int() for _ in range(4)]
        _check(lib().cqx_basis_dims(self._h, *[C.byref(d) for d in dims]))
        self.nShell, self.nBasis, self.maxPrim, self.maxL = (d.value for d in dims)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cqx_basis_destroy(self._h)
            self._h = None

    @classmethod
    def batch(cls, parts):
        """Several molecules' BasisSets side by side (cqx_basis_batch): a
        DirectTPI on the result contracts all molecules in shared launches;
        its twoBodyContract takes the molecules' contractions back to back."""
        arr = (C.c_void_p * len(parts))(*[p._h.value for p in parts])
        h = C.c_void_p()
        _check(lib().cqx_basis_batch(len(parts), C.cast(arr, C.c_void_p), C.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.mol = None
        self.parts = list(parts)
        dims = [C.c_int() for _ in range(4)]
        _check(lib().cqx_basis_dims(self._h, *[C.byref(d) for d in dims]))
        self.nShell, self.nBasis, self.maxPrim, self.maxL = (d.value for d in dims)
        return self

    def one_body(self, kind, Z=None, xyz=None):
        """S, T or V (point nuclei; default: this basis' molecule) as an
        NB x NB array -- OnePInts::computeAOInts (aointegrals_onee_drivers.cxx:762-784)."""
        out = np.zeros((self.nBasis, self.nBasis), order="F")
        if kind == POTENTIAL:
            Z = np.ascontiguousarray(self.mol.Z if Z is None else Z, dtype=np.float64)
            xyz = np.ascontiguousarray(self.mol.xyz if xyz is None else xyz, dtype=np.float64)
            _check(lib().cqx_one_body(self._h, kind, len(Z), _ptr(Z), _ptr(xyz), _ptr(out)))
        else:
            _check(lib().cqx_one_body(self._h, kind, 0, None, None, _ptr(out)))
        return out

    def shell(self, s):
        v = [C.c_int() for _ in range(5)]
        _check(lib().cqx_basis_shell(self._h, s, *[C.byref(x) for x in v]))
        l, pure, nprim, bf, atom = (x.value for x in v)
        a, c, o = np.zeros(nprim), np.zeros(nprim), np.zeros(3)
        _check(lib().cqx_basis_shell_prims(self._h, s, _ptr(a), _ptr(c), _ptr(o)))
        return dict(l=l, pure=pure, nprim=nprim, bf=bf, atom=atom, alpha=a, coef=c, origin=o)


class DirectTPI:
    """Device-resident DirectTPI<double>(basis, basis, mol, threshSchwarz)."""

    def __init__(self, basis, thresh_schwarz=1e-12, device=-1):
        self.basis = basis
        self.threshSchwarz = thresh_schwarz
        h = C.c_void_p()
        _check(lib().cqx_ctx_create(basis._h, float(thresh_schwarz), int(device), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cqx_ctx_destroy(self._h)
            self._h = None

    def schwarz(self):
        ns = self.basis.nShell
        Q = np.zeros((ns, ns), order="F")
        _check(lib().cqx_ctx_schwarz(self._h, _ptr(Q)), self._h)
        return Q

    def shell_pairs(self):
        n = C.c_int()
        _check(lib().cqx_ctx_shell_pairs(self._h, C.byref(n), None, None), self._h)
        s1 = np.zeros(n.value, dtype=np.int32)
        s2 = np.zeros(n.value, dtype=np.int32)
        _check(lib().cqx_ctx_shell_pairs(self._h, C.byref(n), _ptr(s1), _ptr(s2)), self._h)
        return s1, s2

    def eri_quartets(self, quartets):
        q = np.asarray(quartets, dtype=np.int32).reshape(-1, 4)
        cols = [np.ascontiguousarray(q[:, k]) for k in range(4)]
        nf = []
        for s in range(self.basis.nShell):
            sh = self.basis.shell(s)
            nf.append(2 * sh["l"] + 1 if sh["pure"] else (sh["l"] + 1) * (sh["l"] + 2) // 2)
        total = sum(nf[a] * nf[b] * nf[c] * nf[d] for a, b, c, d in q)
        out = np.zeros(max(total, 1))
        off = np.zeros(len(q) + 1, dtype=np.uint64)
        _check(lib().cqx_eri_quartets(self._h, len(q), *[_ptr(c) for c in cols], _ptr(out),
                                      _ptr(off)), self._h)
        return [out[int(off[i]):int(off[i + 1])] for i in range(len(q))]

    def set_option(self, name, value):
        _check(lib().cqx_set_option(self._h, name.encode(), float(value)), self._h)

    def last_stats(self):
        st = Stats()
        _check(lib().cqx_last_stats(self._h, C.byref(st)), self._h)
        return {k: getattr(st, k) for k, _ in Stats._fields_}


class DeviceMatrix:
    """An NB x NB matrix (float64 or complex128, column-major) resident in
    the HBM of a DirectTPI's GPU."""

    def __init__(self, tpi, nb, complex_=False):
        self.tpi, self.nb, self.cplx = tpi, nb, bool(complex_)
        self.nbytes = nb * nb * (16 if complex_ else 8)
        p = C.c_void_p()
        _check(lib().cqx_device_alloc(tpi._h, self.nbytes, C.byref(p)), tpi._h)
        self._p = p

    def data_ptr(self):
        return self._p.value

    def is_complex(self):
        return self.cplx

    def upload(self, M):
        a = np.asfortranarray(M, dtype=np.complex128 if self.cplx else np.float64)
        _check(lib().cqx_memcpy(self.tpi._h, self._p, a.ctypes.data_as(C.c_void_p), self.nbytes, 0),
               self.tpi._h)
        return self

    def download(self):
        a = np.zeros((self.nb, self.nb), dtype=np.complex128 if self.cplx else np.float64, order="F")
        _check(lib().cqx_memcpy(self.tpi._h, a.ctypes.data_as(C.c_void_p), self._p, self.nbytes, 1),
               self.tpi._h)
        return a

    def zero(self):
        _check(lib().cqx_device_zero(self.tpi._h, self._p, self.nbytes), self.tpi._h)
        return self

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None and getattr(self.tpi, "_h", None):
            _lib.cqx_device_free(self.tpi._h, self._p)
            self._p = None


class GTODirectTPIContraction:
    """TPIContractions implementation for the DIRECT algorithm."""

    def __init__(self, tpi):
        if not isinstance(tpi, DirectTPI):
            raise CErr("GTODirectTPIContraction expects a DirectTPI")
        self.tpi = tpi

    def twoBodyContract(self, contractions, screen=True):
        """AX += J[X] / K[X] for every contraction (callers zero AX)."""
        if not contractions:
            raise CErr("Nothing specified to compute")
        dev = [hasattr(c.X, "data_ptr") for c in contractions]
        if any(dev) != all(dev):
            raise CErr("contraction list mixes host and device matrices")
        arr = (_Contraction * len(contractions))()
        if all(dev):
            cplx = contractions[0].X.is_complex()
            for i, c in enumerate(contractions):
                arr[i].X, arr[i].AX = c.X.data_ptr(), c.AX.data_ptr()
                arr[i].her, arr[i].type = int(c.HER), c.contType
            _check(lib().cqx_two_body_contract(self.tpi._h, int(screen), len(contractions),
                                               C.cast(arr, C.c_void_p), int(cplx), DEVICE),
                   self.tpi._h)
            return
        cplx = any(np.iscomplexobj(c.X) or np.iscomplexobj(c.AX) for c in contractions)
        dt = np.complex128 if cplx else np.float64
        keep = []
        for i, c in enumerate(contractions):
            X = np.asfortranarray(c.X, dtype=dt)
            if not (isinstance(c.AX, np.ndarray) and c.AX.dtype == dt and c.AX.flags.f_contiguous):
                raise CErr("AX must be a Fortran-ordered array of the contraction dtype")
            keep.append(X)
            arr[i].X, arr[i].AX = X.ctypes.data, c.AX.ctypes.data
            arr[i].her, arr[i].type = int(c.HER), c.contType
        _check(lib().cqx_two_body_contract(self.tpi._h, int(screen), len(contractions),
                                           C.cast(arr, C.c_void_p), int(cplx), HOST),
               self.tpi._h)


class DirectGradContraction:
    """GradContractions implementation for the DIRECT algorithm
    (particleintegrals/gradints/direct.hpp:31-76): gradTwoBodyContract takes
    cList[g][m], g = 3 atom + xyz, with one density per matrix m shared by
    every g, and adds the derivative-integral J / K of coordinate g to
    cList[g][m].AX (direct.hpp:2198-2580).  Real Hermitian densities."""

    def __init__(self, tpi):
        if not isinstance(tpi, DirectTPI):
            raise CErr("DirectGradContraction expects a DirectTPI")
        self.tpi = tpi

    def gradTwoBodyContract(self, cList, screen=True):
        ngrad, nmat = len(cList), len(cList[0]) if cList else 0
        if not ngrad or not nmat or any(len(x) != nmat for x in cList):
            raise CErr("gradTwoBodyContract takes the same number of matrices per coordinate")
        arr = (_Contraction * (ngrad * nmat))()
        X = [np.asfortranarray(c.X, dtype=np.float64) for c in cList[0]]
        for g, row in enumerate(cList):
            for m, c in enumerate(row):
                if not (isinstance(c.AX, np.ndarray) and c.AX.dtype == np.float64 and c.AX.flags.f_contiguous):
                    raise CErr("AX must be a Fortran-ordered float64 array")
                if c.X is not cList[0][m].X:
                    raise CErr("every coordinate must contract the same density per matrix")
                k = g * nmat + m
                arr[k].X, arr[k].AX = X[m].ctypes.data, c.AX.ctypes.data
                arr[k].her, arr[k].type = int(c.HER), c.contType
        _check(lib().cqx_grad_two_body_contract(self.tpi._h, ngrad, nmat, C.cast(arr, C.c_void_p), HOST),
               self.tpi._h)


def getGDGrad(onePDM, tpi, natom, xHFX=1.0):
    """FockBuilder::getGDGrad (fockbuilder/impl.hpp:371-488) for a real
    closed-shell density: J^x[D_S] and K^x[D_S] per nuclear coordinate,
    G^x = 2 J^x - xHFX K^x, and the two-electron gradient
    1/4 Tr(D_S G^x) (computeOBProperty<SCALAR>)."""
    D = np.asfortranarray(onePDM, dtype=np.float64)
    nb = D.shape[0]
    ngrad = 3 * natom
    doK = abs(xHFX) > 1e-12
    J = [np.zeros((nb, nb), order="F") for _ in range(ngrad)]
    K = [np.zeros((nb, nb), order="F") for _ in range(ngrad)] if doK else None
    cList = []
    for g in range(ngrad):
        row = [TwoBodyContraction(D, J[g], True, COULOMB)]
        if doK:
            row.append(TwoBodyContraction(D, K[g], True, EXCHANGE))
        cList.append(row)
    DirectGradContraction(tpi).gradTwoBodyContract(cList)
    grad = np.zeros(ngrad)
    for g in range(ngrad):
        G = 2.0 * J[g] - (xHFX * K[g] if doK else 0.0)
        grad[g] = 0.25 * float(np.sum(D * G))
    return grad, J, K


class SingleSlaterState:
    """The pieces of SingleSlater that FockBuilder::formGD reads and writes:
    onePDM as Pauli components [S] (1C) or [S, Z, Y, X] (2C), and the
    outputs coulombMatrix, exchangeMatrix[c], twoeH[c]."""

    def __init__(self, onePDM, TPI):
        self.onePDM = [np.asfortranarray(d) for d in onePDM]
        self.TPI = TPI
        nb = self.onePDM[0].shape[0]
        dt = np.result_type(*[d.dtype for d in self.onePDM])
        self.coulombMatrix = np.zeros((nb, nb), dtype=dt, order="F")
        self.exchangeMatrix = [np.zeros((nb, nb), dtype=dt, order="F") for _ in self.onePDM]
        self.twoeH = [np.zeros((nb, nb), dtype=dt, order="F") for _ in self.onePDM]


class FockBuilder:
    """FockBuilder<MatsT,IntsT>::formGD (include/fockbuilder/impl.hpp:75-244)
    for electronic, non-incremental builds: J on S, K on S (+ Z, Y, X for
    two-component densities), G = 2J - xHFX K componentwise."""

    @staticmethod
    def formGD(ss, xHFX=1.0, HerDen=True):
        dt = ss.coulombMatrix.dtype
        D = [np.asfortranarray(d, dtype=dt) for d in ss.onePDM]
        J = np.zeros_like(ss.coulombMatrix, order="F")
        K = [np.zeros_like(ss.coulombMatrix, order="F") for _ in D]
        cont = [TwoBodyContraction(D[0], J, HerDen, COULOMB)]
        doK = abs(xHFX) > 1e-12
        if doK:
            cont += [TwoBodyContraction(D[c], K[c], HerDen, EXCHANGE) for c in range(len(D))]
        ss.TPI.twoBodyContract(cont)
        ss.coulombMatrix[...] = J
        for c in range(len(D)):
            ss.exchangeMatrix[c][...] = K[c]
            ss.twoeH[c][...] = (2.0 * J if c == 0 else 0.0) - (xHFX * K[c] if doK else 0.0)
        return ss


class ScfOpts(C.Structure):
    _fields_ = [("nocc", C.c_int), ("max_iter", C.c_int), ("rmsd_tol", C.c_double),
                ("maxd_tol", C.c_double), ("ene_tol", C.c_double), ("diis_keep", C.c_int),
                ("guess", C.c_int), ("fp32_cut", C.c_double)]


class ScfResult(C.Structure):
    _fields_ = [("e_total", C.c_double), ("e_one", C.c_double), ("e_two", C.c_double),
                ("e_nuc", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("rmsd", C.c_double), ("maxd", C.c_double), ("delta_e", C.c_double),
                ("ms_setup", C.c_double), ("ms_fock", C.c_double), ("ms_total", C.c_double)]


GUESS = {"CORE": 0, "SAD": 1}
SCF_DENSITY, SCF_FOCK, SCF_HCORE, SCF_OVERLAP = 0, 1, 2, 3


def _scf_lib():
    L = lib()
    if not getattr(L, "_scf_typed", False):
        P, I = C.c_void_p, C.c_int
        L.cqx_scf_default_opts.argtypes = [C.POINTER(ScfOpts)]
        L.cqx_scf_default_opts.restype = None
        L.cqx_scf_create.argtypes = [P, I, P, P, C.POINTER(ScfOpts), C.POINTER(P)]
        L.cqx_scf_iterate.argtypes = [P, C.POINTER(I)]
        L.cqx_scf_run.argtypes = [P, C.POINTER(ScfResult)]
        L.cqx_scf_result_get.argtypes = [P, C.POINTER(ScfResult)]
        L.cqx_scf_matrix.argtypes = [P, I, P]
        L.cqx_scf_orbital_energies.argtypes = [P, P]
        L.cqx_scf_destroy.argtypes = [P]
        L.cqx_scf_last_error.argtypes = [P]
        L.cqx_scf_last_error.restype = C.c_char_p
        L._scf_typed = True
    return L


class RHF:
    """Closed-shell real Hartree-Fock: SingleSlater<double,double> with the
    conventional SCF (orbitaloptimizernew/impl.hpp:40-137), run on the GPU
    of a DirectTPI -- one-electron integrals, Lowdin orthogonaliser, SAD or
    core guess, CDIIS, eigensolve and density all device-resident; Fock
    builds through the direct J/K engine.  Energies in Hartree, matrices in
    the Pauli-scalar convention (D = 2 C C^T, H = 2 (T + V))."""

    def __init__(self, tpi, nocc=None, guess="SAD", fp32_cut=0.0, max_iter=128,
                 rmsd_tol=1e-8, maxd_tol=1e-5, ene_tol=1e-5, diis_keep=8):
        L = _scf_lib()
        mol = tpi.basis.mol
        if nocc is None:
            ne = mol.nelectron()
            if ne % 2:
                raise CErr("RHF needs an even electron count")
            nocc = ne // 2
        o = ScfOpts()
        L.cqx_scf_default_opts(C.byref(o))
        o.nocc, o.max_iter, o.rmsd_tol, o.maxd_tol = int(nocc), int(max_iter), rmsd_tol, maxd_tol
        o.ene_tol, o.diis_keep, o.guess, o.fp32_cut = ene_tol, int(diis_keep), GUESS[guess.upper()], fp32_cut
        self.tpi, self.nb = tpi, tpi.basis.nBasis
        self._Z = np.ascontiguousarray(mol.Z, dtype=np.float64)
        self._xyz = np.ascontiguousarray(mol.xyz, dtype=np.float64)
        h = C.c_void_p()
        rc = L.cqx_scf_create(tpi._h, len(self._Z), _ptr(self._Z), _ptr(self._xyz), C.byref(o),
                              C.byref(h))
        if rc:
            raise CErr(f"FATAL (cqx status {rc}): {L.cqx_scf_last_error(None).decode()}")
        self._h = h

    def _chk(self, rc):
        if rc:
            raise CErr(f"FATAL (cqx status {rc}): {_scf_lib().cqx_scf_last_error(self._h).decode()}")

    def iterate(self):
        done = C.c_int()
        self._chk(_scf_lib().cqx_scf_iterate(self._h, C.byref(done)))
        return bool(done.value)

    def result(self):
        r = ScfResult()
        self._chk(_scf_lib().cqx_scf_result_get(self._h, C.byref(r)))
        return {k: getattr(r, k) for k, _ in ScfResult._fields_}

    def SCF(self):
        """Iterate to convergence (CErr after max_iter, like the reference)."""
        r = ScfResult()
        self._chk(_scf_lib().cqx_scf_run(self._h, C.byref(r)))
        return {k: getattr(r, k) for k, _ in ScfResult._fields_}

    def matrix(self, which):
        out = np.zeros((self.nb, self.nb), order="F")
        self._chk(_scf_lib().cqx_scf_matrix(self._h, which, _ptr(out)))
        return out

    def orbital_energies(self):
        out = np.zeros(self.nb)
        self._chk(_scf_lib().cqx_scf_orbital_energies(self._h, _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cqx_scf_destroy(self._h)
            self._h = None


class _Job(C.Structure):
    _fields_ = [("g94", C.c_char_p), ("natom", C.c_int), ("Z", C.c_void_p), ("xyz", C.c_void_p),
                ("nocc", C.c_int), ("arrival_s", C.c_double)]


class ServeOpts(C.Structure):
    _fields_ = [("max_inflight", C.c_int), ("hbm_budget_bytes", C.c_double), ("device", C.c_int),
                ("thresh", C.c_double), ("list_cap", C.c_longlong), ("scf", ScfOpts),
                ("verbose", C.c_int), ("graph", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int), ("converged", C.c_int), ("iterations", C.c_int),
                ("e_total", C.c_double), ("t_arrival", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("footprint_bytes", C.c_double), ("ms_fock", C.c_double)]


def serve(jobs, max_inflight=8, hbm_budget_bytes=200e9, fp32_cut=0.0, guess="SAD",
          list_cap=1 << 25, device=0, rmsd_tol=1e-8, verbose=0, graph=0):
    """Run SCF jobs [(basis name or G94 text, Molecule, arrival seconds)]
    through the serving loop; returns one dict per job (latency =
    t_done - t_arrival)."""
    L = _scf_lib()
    if not getattr(L, "_serve_typed", False):
        L.cqx_serve_default_opts.argtypes = [C.POINTER(ServeOpts)]
        L.cqx_serve_default_opts.restype = None
        L.cqx_serve.argtypes = [C.c_void_p, C.c_int, C.POINTER(ServeOpts), C.c_void_p]
        L._serve_typed = True
    o = ServeOpts()
    L.cqx_serve_default_opts(C.byref(o))
    o.max_inflight, o.hbm_budget_bytes, o.device, o.list_cap = int(max_inflight), float(hbm_budget_bytes), int(device), int(list_cap)
    o.verbose, o.graph = int(verbose), int(graph)
    o.scf.fp32_cut, o.scf.guess, o.scf.rmsd_tol = fp32_cut, GUESS[guess.upper()], rmsd_tol
    arr = (_Job * len(jobs))()
    keep = []
    for i, (bas, mol, t) in enumerate(jobs):
        text = (bas if "****" in bas else basis_text(bas)).encode()
        Z = np.ascontiguousarray(mol.Z, dtype=np.int32)
        xyz = np.ascontiguousarray(mol.xyz, dtype=np.float64)
        keep += [text, Z, xyz]
        ne = mol.nelectron()
        if ne % 2:
            raise CErr("RHF serving jobs need an even electron count")
        arr[i].g94, arr[i].natom, arr[i].Z, arr[i].xyz = text, len(Z), Z.ctypes.data, xyz.ctypes.data
        arr[i].nocc, arr[i].arrival_s = ne // 2, float(t)
    res = (JobResult * len(jobs))()
    rc = L.cqx_serve(C.cast(arr, C.c_void_p), len(jobs), C.byref(o), C.cast(res, C.c_void_p))
    out = [{k: getattr(r, k) for k, _ in JobResult._fields_} for r in res]
    if rc:
        bad = [i for i, r in enumerate(out) if r["status"]]
        raise CErr(f"FATAL (cqx status {rc}): serving jobs {bad} failed")
    return out

