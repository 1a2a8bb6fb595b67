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
# CQX_NOALLOC_SCRATCH_BYTES: the fixed per-GPU workspace region of the
# allocation-free contraction
NOALLOC_SCRATCH_BYTES = 64 << 20


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
                ("flops_model_f32", C.c_double), ("graph", C.c_int),
                ("workspace_bytes", C.c_longlong)]


SYMBOLS = ["cqx_basis_create", "cqx_basis_from_g94", "cqx_basis_destroy",
           "cqx_basis_dims", "cqx_basis_shell", "cqx_basis_shell_prims",
           "cqx_ctx_create", "cqx_ctx_destroy", "cqx_last_error", "cqx_ctx_schwarz",
           "cqx_ctx_shell_pairs", "cqx_two_body_contract", "cqx_eri_quartets",
           "cqx_last_stats", "cqx_set_option", "cqx_device_alloc", "cqx_device_free",
           "cqx_memcpy", "cqx_device_zero", "cqx_one_body", "cqx_basis_batch", "cqx_scf_default_opts",
           "cqx_scf_create", "cqx_scf_iterate", "cqx_scf_run", "cqx_scf_result_get",
           "cqx_scf_matrix", "cqx_scf_orbital_energies", "cqx_scf_destroy",
           "cqx_scf_last_error", "cqx_serve_default_opts", "cqx_serve", "cqx_grad_two_body_contract",
           "cqx_scratch_size", "cqx_spin_scatter", "cqx_spin_gather", "cqx_compute_energy",
           "cqx_count_nonfinite", "cqx_two_body_contract_noalloc"]

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
        L.cqx_scratch_size.argtypes = [P, I, P, I, C.POINTER(C.c_size_t)]
        L.cqx_two_body_contract_noalloc.argtypes = [P, I, I, P, I]
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

    def scratch_size(self, types, hers=None, is_complex=False):
        """GTODirectTPIContraction::directScaffoldNewSCRSize (direct.hpp:2125-2180):
        device bytes one contraction call with these types (and HER flags)
        allocates at most (cqx_scratch_size)."""
        types = list(types)
        hers = [True] * len(types) if hers is None else list(hers)
        arr = (_Contraction * len(types))()
        for i, (t, h) in enumerate(zip(types, hers)):
            arr[i].X, arr[i].AX, arr[i].her, arr[i].type = None, None, int(h), int(t)
        b = C.c_size_t()
        _check(lib().cqx_scratch_size(self._h, len(types), C.cast(arr, C.c_void_p), int(is_complex),
                                      C.byref(b)), self._h)
        return b.value

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

    def twoBodyContract(self, *args, screen=True, pert=None):
        """AX += J[X] / K[X] for every contraction (callers zero AX).

        The reference's forms (twopints.hpp:219-244) -- (comm, screen,
        contList, pert), (comm, screen, contList), (comm, contList, pert),
        (comm, contList) -- and the short (contList[, screen]).  The
        perturbation is not read, as GTODirectTPIContraction ignores it
        (gtodirecttpi.hpp:199-206); comm is the single rank of this process."""
        pos = [i for i, a in enumerate(args) if isinstance(a, (list, tuple))]
        if not pos:
            raise CErr("twoBodyContract: no contraction list")
        i = pos[0]
        contractions = args[i]
        if i == 0 and len(args) > 1:
            screen = bool(args[1])
        elif i == 2:
            screen = bool(args[1])
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

    def twoBodyContractNoAlloc(self, contractions, screen=True):
        """The same contraction through cqx_two_body_contract_noalloc: X and
        AX are DeviceMatrix objects, AX is accumulated in place, and every
        per-call buffer comes from the library's fixed scratch region
        (no device allocation during the call)."""
        if not contractions:
            raise CErr("Nothing specified to compute")
        if not all(hasattr(c.X, "data_ptr") and hasattr(c.AX, "data_ptr") for c in contractions):
            raise CErr("the allocation-free contraction takes device matrices")
        arr = (_Contraction * len(contractions))()
        for i, c in enumerate(contractions):
            arr[i].X, arr[i].AX = c.X.data_ptr(), c.AX.data_ptr()
            arr[i].her, arr[i].type = int(c.HER), c.contType
        _check(lib().cqx_two_body_contract_noalloc(self.tpi._h, int(screen), len(contractions),
                                                   C.cast(arr, C.c_void_p),
                                                   int(contractions[0].X.is_complex())),
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


SPIN_ZERO_ABBA, SPIN_BB_EQ_AA, SPIN_ZERO_XY, SPIN_ZERO_Z = 1, 2, 4, 8


def _pauli_lib():
    L = lib()
    if not getattr(L, "_pauli_typed", False):
        P, I = C.c_void_p, C.c_int
        L.cqx_spin_scatter.argtypes = [P, I, P, I, P, P, P, P, I, I, I]
        L.cqx_spin_gather.argtypes = [P, I, P, I, P, P, P, P, I, I, I]
        L.cqx_compute_energy.argtypes = [P, I, I, P, P, P, I, I, C.c_double, P]
        L.cqx_count_nonfinite.argtypes = [P, P, C.c_size_t, I, C.POINTER(C.c_size_t)]
        L._pauli_typed = True
    return L


def _tpi_handle(tpi):
    """The context handle behind a DirectTPI or a contraction object."""
    t = tpi.tpi if isinstance(tpi, GTODirectTPIContraction) else tpi
    return t._h


def SpinScatter(tpi, A, zeroABBA=False, BBeqAA=False):
    """SpinScatter (src/cqlinalg/blasutil.cxx:71-129) on the GPU of a
    DirectTPI: the 2NB x 2NB spinor matrix A -> [S, Z, Y, X], NB x NB each,
    S = AA + BB, Z = AA - BB, Y = f (AB - BA), X = AB + BA (f = i complex,
    -1 real)."""
    A = np.asfortranarray(A)
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    A = A.astype(dt, order="F", copy=False)
    nb = A.shape[0] // 2
    if A.shape != (2 * nb, 2 * nb):
        raise CErr("SpinScatter takes a square matrix of even dimension")
    out = [np.zeros((nb, nb), dtype=dt, order="F") for _ in range(4)]
    flags = (SPIN_ZERO_ABBA if zeroABBA else 0) | (SPIN_BB_EQ_AA if BBeqAA else 0)
    h = _tpi_handle(tpi)
    _check(_pauli_lib().cqx_spin_scatter(h, nb, _ptr(A), 2 * nb, *[_ptr(o) for o in out], int(cplx), flags, HOST), h)
    return out


def SpinGather(tpi, S, Z=None, Y=None, X=None):
    """SpinGather (blasutil.cxx:131-189): [S, Z, Y, X] -> the 2NB x 2NB
    spinor matrix (Z None: zeroZ; Y and X None: zeroXY)."""
    comps = [S, Z, Y, X]
    cplx = any(c is not None and np.iscomplexobj(c) for c in comps)
    dt = np.complex128 if cplx else np.float64
    arrs = [None if c is None else np.asfortranarray(c, dtype=dt) for c in comps]
    nb = arrs[0].shape[0]
    zero_z, zero_xy = arrs[1] is None, arrs[2] is None or arrs[3] is None
    flags = (SPIN_ZERO_Z if zero_z else 0) | (SPIN_ZERO_XY if zero_xy else 0)
    A = np.zeros((2 * nb, 2 * nb), dtype=dt, order="F")
    ptrs = [None if a is None else _ptr(a) for a in arrs]
    h = _tpi_handle(tpi)
    _check(_pauli_lib().cqx_spin_gather(h, nb, _ptr(A), 2 * nb, *ptrs, int(cplx), flags, HOST), h)
    return A


class PauliSpinorMatrices:
    """cqmatrix::PauliSpinorMatrices (include/matrix/paulispinormatrices.hpp:30-392):
    the S component and, where present, Z (hasZ) and Y, X (hasXY); each an
    NB x NB column-major array."""

    def __init__(self, nb, hasXY=False, hasZ=False, dtype=np.float64):
        self.hasZ, self.hasXY = bool(hasZ or hasXY), bool(hasXY)
        self.dtype = np.dtype(dtype)
        self.comps = [np.zeros((nb, nb), dtype=dtype, order="F") for _ in range(1 + self.hasZ + 2 * self.hasXY)]

    @classmethod
    def from_components(cls, comps):
        comps = [np.asfortranarray(c) for c in comps]
        dt = np.result_type(*[c.dtype for c in comps])
        self = cls(comps[0].shape[0], hasXY=len(comps) == 4, hasZ=len(comps) >= 2, dtype=dt)
        for k, c in enumerate(comps):
            self.comps[k][...] = c
        return self

    @classmethod
    def spinScatter(cls, tpi, A, zeroABBA=False, BBeqAA=False):
        return cls.from_components(SpinScatter(tpi, A, zeroABBA, BBeqAA))

    def spinGather(self, tpi):
        return SpinGather(tpi, *(self.comps + [None] * (4 - len(self.comps))))

    def S(self):
        return self.comps[0]

    def Z(self):
        if not self.hasZ:
            raise CErr("PauliSpinorMatrices has no Z component")
        return self.comps[1]

    def Y(self):
        if not self.hasXY:
            raise CErr("PauliSpinorMatrices has no Y component")
        return self.comps[2]

    def X(self):
        if not self.hasXY:
            raise CErr("PauliSpinorMatrices has no X component")
        return self.comps[3]

    def clear(self):
        for c in self.comps:
            c[...] = 0

    def hasNaN(self, tpi):
        """PauliSpinorMatrices::hasNaN (paulispinormatrices.hpp:383-388), on the GPU."""
        h = _tpi_handle(tpi)
        n = C.c_size_t()
        for c in self.comps:
            flat = np.ascontiguousarray(c).view(np.float64)
            _check(_pauli_lib().cqx_count_nonfinite(h, _ptr(flat), flat.size, HOST, C.byref(n)), h)
            if n.value:
                return True
        return False


def computeEnergy(tpi, onePDM, coreH, twoeH, nucRepEnergy=0.0):
    """SingleSlater::computeEnergy (include/singleslater/quantum.hpp:147-212)
    on the GPU: (OBEnergy, MBEnergy, totalEnergy) = (1/2 sum_c Re<D_c,H_c>,
    1/4 sum_c Re<D_c,G_c>, their sum + V_NN) over the Pauli components."""
    D = [np.asfortranarray(d) for d in onePDM]
    cplx = any(np.iscomplexobj(x) for x in list(D) + list(coreH or []) + list(twoeH or []))
    dt = np.complex128 if cplx else np.float64
    D = [d.astype(dt, order="F", copy=False) for d in D]
    nc, nb = len(D), D[0].shape[0]

    def arr(mats):
        if mats is None:
            return None, []
        keep = [None if m is None else np.asfortranarray(m, dtype=dt) for m in list(mats)[:nc]]
        keep += [None] * (nc - len(keep))
        return (C.c_void_p * nc)(*[None if k is None else k.ctypes.data for k in keep]), keep

    pd, kd = arr(D)
    ph, kh = arr(coreH)
    pg, kg = arr(twoeH)
    out = np.zeros(3)
    h = _tpi_handle(tpi)
    _check(_pauli_lib().cqx_compute_energy(h, nb, nc, C.cast(pd, C.c_void_p),
                                          None if ph is None else C.cast(ph, C.c_void_p),
                                          None if pg is None else C.cast(pg, C.c_void_p),
                                          int(cplx), HOST, float(nucRepEnergy), _ptr(out)), h)
    return float(out[0]), float(out[1]), float(out[2])


class EMPerturbation:
    """EMPerturbation (include/fields.hpp:124-195), dipole fields only:
    fields is a list of (type, amplitude[3]) with type "Electric" or
    "Magnetic"."""

    def __init__(self, fields=None):
        self.fields = list(fields or [])

    def addField(self, typ, amp):
        self.fields.append((typ, np.asarray(amp, dtype=float)))

    def has(self, typ):
        return any(t == typ for t, _ in self.fields)

    def getDipoleAmp(self, typ):
        amp = np.zeros(3)
        for t, a in self.fields:
            if t == typ:
                amp += a
        return amp


class SingleSlaterState:
    """The pieces of SingleSlater that FockBuilder reads and writes: onePDM
    (and deltaOnePDM for incremental builds) as Pauli components [S] (1C) or
    [S, Z, Y, X] (2C); the outputs coulombMatrix, exchangeMatrix[c],
    twoeH[c], fockMatrix[c]; coreH[c] for formFock; nC; the contraction
    engine TPI; GDDur (seconds of the last formGD inside formFock)."""

    def __init__(self, onePDM, TPI, coreH=None, deltaOnePDM=None, nC=None, lenElectric=None):
        self.onePDM = [np.asfortranarray(d) for d in onePDM]
        self.TPI = TPI
        nb = self.onePDM[0].shape[0]
        mats = list(self.onePDM) + list(coreH or []) + list(deltaOnePDM or [])
        dt = np.result_type(*[m.dtype for m in mats])
        self.deltaOnePDM = None if deltaOnePDM is None else [np.asfortranarray(d, dtype=dt) for d in deltaOnePDM]
        self.coreH = None if coreH is None else [np.asfortranarray(h, dtype=dt) for h in coreH]
        self.nC = nC if nC is not None else (2 if len(self.onePDM) > 1 else 1)
        self.lenElectric = lenElectric
        self.coulombMatrix = np.zeros((nb, nb), dtype=dt, order="F")
        self.exchangeMatrix = [np.zeros((nb, nb), dtype=dt, order="F") for _ in self.onePDM]
        self.twoeH = [np.zeros((nb, nb), dtype=dt, order="F") for _ in self.onePDM]
        self.fockMatrix = [np.zeros((nb, nb), dtype=dt, order="F") for _ in self.onePDM]
        self.GDDur = 0.0


class FockBuilder:
    """FockBuilder<MatsT,IntsT> (include/fockbuilder/impl.hpp:75-319) for
    electronic densities: formGD, formRawGDInBatches, formRawGDSCRSizePerBatch
    and formFock, with the reference's list building (J on S; K on S, Z and,
    for two-component densities, Y and X), zeroing unless incremental, and
    G = 2J - xHFX K componentwise."""

    @staticmethod
    def formGD(ss, pert=None, increment=False, xHFX=1.0, HerDen=True):
        """impl.hpp:75-102: G[D]; with increment, the reference's
        accumulation on deltaOnePDM: coulombMatrix = J[delta],
        K += K[delta] and G += 2 J[delta] - xHFX K, with K the accumulated
        exchange matrix (impl.hpp:223-237)."""
        src = ss.deltaOnePDM if increment else ss.onePDM
        if src is None:
            raise CErr("formGD: increment needs deltaOnePDM")
        dt = ss.coulombMatrix.dtype
        J = [np.zeros_like(ss.coulombMatrix, order="F")]
        FockBuilder.formRawGDInBatches(ss, pert, increment, xHFX, HerDen, [src], J,
                                       [ss.exchangeMatrix], [ss.twoeH])
        ss.coulombMatrix[...] = J[0].astype(dt, copy=False)
        return ss

    @staticmethod
    def formRawGDInBatches(ss, pert, increment, xHFX, HerDen, onePDMs, coulombMatrices,
                           exchangeMatrices, twoeHs):
        """impl.hpp:109-244: every batch b takes onePDMs[b] (Pauli
        components) and writes coulombMatrices[b] (the S component, one
        NB x NB array), exchangeMatrices[b][c] and twoeHs[b][c]; all batches
        go to one twoBodyContract call."""
        nBatch = len(onePDMs)
        computeCoulomb = coulombMatrices is not None and len(coulombMatrices) > 0
        computeExchange = abs(xHFX) > 1e-12 and exchangeMatrices is not None and len(exchangeMatrices) > 0
        computeTwoeHs = twoeHs is not None and len(twoeHs) > 0
        if not (computeCoulomb or computeExchange or computeTwoeHs):
            raise CErr("Nothing specified to compute in FockBuilder::formRawGDInBatches")
        if getattr(ss, "nC", 1) == 4:
            raise CErr("4C formGD is implemented in class FourCompFock.")
        if not increment:
            for i in range(nBatch):
                if computeCoulomb:
                    coulombMatrices[i][...] = 0
                if computeExchange:
                    for m in exchangeMatrices[i]:
                        m[...] = 0
                if computeTwoeHs:
                    for m in twoeHs[i]:
                        m[...] = 0
        dt = np.result_type(*[np.asarray(d).dtype for b in onePDMs for d in b])
        if computeCoulomb:
            dt = np.result_type(dt, coulombMatrices[0].dtype)
        D = [[np.asfortranarray(d, dtype=dt) for d in b] for b in onePDMs]
        cont = []
        if computeCoulomb or computeTwoeHs:
            cont_j = coulombMatrices if computeCoulomb else [t[0] for t in twoeHs]
            for i in range(nBatch):
                cont.append(TwoBodyContraction(D[i][0], cont_j[i], HerDen, COULOMB))
        if computeExchange:
            for i in range(nBatch):
                for c in range(len(exchangeMatrices[i])):
                    cont.append(TwoBodyContraction(D[i][c], exchangeMatrices[i][c], HerDen, EXCHANGE))
        ss.TPI.twoBodyContract(cont)
        if computeTwoeHs:
            for i in range(nBatch):
                if computeCoulomb:
                    twoeHs[i][0] += 2.0 * coulombMatrices[i]
                else:
                    twoeHs[i][0] *= 2.0
                if computeExchange:
                    for c in range(len(twoeHs[i])):
                        twoeHs[i][c] -= xHFX * exchangeMatrices[i][c]
        return ss

    @staticmethod
    def formRawGDSCRSizePerBatch(ss, computeExchange=True, HerDen=True):
        """impl.hpp:252-269: device workspace (in MatsT elements) of one
        batch's contractions (J, plus K on every Pauli component)."""
        cplx = np.iscomplexobj(ss.coulombMatrix)
        nK = len(ss.onePDM) if computeExchange else 0
        types = [COULOMB] + [EXCHANGE] * nK
        b = ss.TPI.tpi.scratch_size(types, [HerDen] * len(types), cplx)
        return -(-b // (16 if cplx else 8))

    @staticmethod
    def formFock(ss, pert=None, increment=False, xHFX=1.0):
        """impl.hpp:280-319: G[D] by formGD (timed into GDDur), then
        F = H + G componentwise, minus 2 E.mu_i for an electric dipole field
        (ss.lenElectric: the three length-gauge dipole matrices)."""
        import time
        if ss.coreH is None:
            raise CErr("formFock needs the core Hamiltonian")
        t0 = time.perf_counter()
        FockBuilder.formGD(ss, pert, increment, xHFX)
        ss.GDDur = time.perf_counter() - t0
        for c in range(len(ss.fockMatrix)):
            ss.fockMatrix[c][...] = ss.coreH[c] + ss.twoeH[c]
        if pert is not None and pert.has("Electric"):
            if ss.lenElectric is None:
                raise CErr("formFock: an electric field needs the length-gauge dipole integrals")
            amp = pert.getDipoleAmp("Electric")
            for i in range(3):
                ss.fockMatrix[0][...] -= 2.0 * amp[i] * ss.lenElectric[i]
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
                ("nocc", C.c_int), ("arrival_s", C.c_double), ("density_out", C.c_void_p),
                ("fock_out", C.c_void_p)]


class ServeOpts(C.Structure):
    _fields_ = [("max_inflight", C.c_int), ("hbm_budget_bytes", C.c_double), ("device", C.c_int),
                ("thresh", C.c_double), ("list_cap", C.c_longlong), ("scf", ScfOpts),
                ("verbose", C.c_int), ("graph", C.c_int), ("batch_fock", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int), ("converged", C.c_int), ("iterations", C.c_int),
                ("e_total", C.c_double), ("t_arrival", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("footprint_bytes", C.c_double), ("ms_fock", C.c_double)]


def serve(jobs, max_inflight=8, hbm_budget_bytes=200e9, fp32_cut=0.0, guess="SAD",
          list_cap=1 << 25, device=0, rmsd_tol=1e-8, verbose=0, graph=0, matrices=False,
          batch_fock=0, allow_refused=False):
    """Run SCF jobs [(basis name or G94 text, Molecule, arrival seconds)]
    through the serving loop; returns one dict per job (latency =
    t_done - t_arrival), with its final density "D" and Fock matrix "F"
    (Pauli-scalar convention, NB x NB) when matrices is set.  batch_fock:
    the Fock builds of every job in flight go into one batched contraction
    per round (cross-request batching).  allow_refused: jobs refused for
    their footprint (status 3) are reported instead of raised."""
    L = _scf_lib()
    if not getattr(L, "_serve_typed", False):
        L.cqx_serve_default_opts.argtypes = [C.POINTER(ServeOpts)]
        L.cqx_serve_default_opts.restype = None
        L.cqx_serve.argtypes = [C.c_void_p, C.c_int, C.POINTER(ServeOpts), C.c_void_p]
        L._serve_typed = True
    o = ServeOpts()
    L.cqx_serve_default_opts(C.byref(o))
    o.max_inflight, o.hbm_budget_bytes, o.device, o.list_cap = int(max_inflight), float(hbm_budget_bytes), int(device), int(list_cap)
    o.verbose, o.graph, o.batch_fock = int(verbose), int(graph), int(batch_fock)
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
    mats = []
    if matrices:
        for i, (bas, mol, t) in enumerate(jobs):
            nb = BasisSet(bas, mol).nBasis
            D, F = np.zeros((nb, nb), order="F"), np.zeros((nb, nb), order="F")
            mats.append((D, F))
            arr[i].density_out, arr[i].fock_out = D.ctypes.data, F.ctypes.data
    res = (JobResult * len(jobs))()
    rc = L.cqx_serve(C.cast(arr, C.c_void_p), len(jobs), C.byref(o), C.cast(res, C.c_void_p))
    out = [{k: getattr(r, k) for k, _ in JobResult._fields_} for r in res]
    for r, m in zip(out, mats):
        r["D"], r["F"] = m
    bad = [i for i, r in enumerate(out) if r["status"] and not (allow_refused and r["status"] == 3)]
    if bad:
        raise CErr(f"FATAL (cqx status {rc}): serving jobs {bad} failed")
    return out

