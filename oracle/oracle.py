"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.  The C source
(oracle.c) restates ChronusQ's direct J/K scaffold
(include/particleintegrals/contract/direct.hpp:1412-2123); see oracle.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

COULOMB, EXCHANGE = 0, 1


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.oracle_basis_create.restype = P
        L.oracle_basis_create.argtypes = [C.c_char_p, I, P, P, I]
        L.oracle_basis_destroy.argtypes = [P]
        L.oracle_last_error.restype = C.c_char_p
        for f in ("oracle_nbasis", "oracle_nshell", "oracle_maxprim", "oracle_maxl"):
            getattr(L, f).argtypes = [P]
            getattr(L, f).restype = I
        L.oracle_shell_info.argtypes = [P, I, P, P, P, P, P]
        L.oracle_shell_prims.argtypes = [P, I, P, P, P]
        L.oracle_boys.argtypes = [I, D, P]
        L.oracle_overlap.argtypes = [P, P]
        L.oracle_eri_quartet.argtypes = [P, I, I, I, I, P]
        L.oracle_eri_quartet.restype = I
        L.oracle_eri_full.argtypes = [P, P]
        L.oracle_shell_pairs.argtypes = [P, P, P]
        L.oracle_shell_pairs.restype = I
        L.oracle_schwarz.argtypes = [P, P]
        L.oracle_two_body_contract.argtypes = [P, D, I, I, P, I, I]
        L.oracle_two_body_contract.restype = I
        L.oracle_grad_contract.argtypes = [P, I, P, I]
        L.oracle_grad_contract.restype = I
        L.oracle_last_stats.argtypes = [P, P, P]
        L.oracle_prepare.argtypes = [P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def boys(mmax, T):
    F = np.zeros(mmax + 1)
    lib().oracle_boys(mmax, float(T), _p(F))
    return F


class _Contraction(C.Structure):
    _fields_ = [("X", C.c_void_p), ("AX", C.c_void_p), ("her", C.c_int),
                ("type", C.c_int)]


class OracleBasis:
    """Basis on a molecule: Z (ints), xyz (natom x 3, bohr), G94 text."""

    def __init__(self, gbs_text, Z, xyz_bohr, force_cart=False):
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        h = lib().oracle_basis_create(gbs_text.encode(), len(self.Z), _p(self.Z),
                                      _p(self.xyz), int(force_cart))
        if not h:
            raise RuntimeError(lib().oracle_last_error().decode())
        self._h = C.c_void_p(h)
        self.nbf = lib().oracle_nbasis(self._h)
        self.nshell = lib().oracle_nshell(self._h)
        self.maxprim = lib().oracle_maxprim(self._h)
        self.maxl = lib().oracle_maxl(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.oracle_basis_destroy(self._h)
            self._h = None

    def shells(self):
        out = []
        for s in range(self.nshell):
            v = [C.c_int() for _ in range(5)]
            lib().oracle_shell_info(self._h, s, *[C.byref(x) for x in v])
            l, pure, nprim, bf, atom = (x.value for x in v)
            a, c, o = np.zeros(nprim), np.zeros(nprim), np.zeros(3)
            lib().oracle_shell_prims(self._h, s, _p(a), _p(c), _p(o))
            out.append(dict(l=l, pure=pure, nprim=nprim, bf=bf, atom=atom,
                            alpha=a, coef=c, origin=o))
        return out

    def overlap(self):
        S = np.zeros((self.nbf, self.nbf), order="F")
        lib().oracle_overlap(self._h, _p(S))
        return S

    def eri_quartet(self, s1, s2, s3, s4):
        buf = np.zeros(28 ** 4 // 16 + 20000)
        n = lib().oracle_eri_quartet(self._h, s1, s2, s3, s4, _p(buf))
        return buf[:n].copy()

    def eri_full(self):
        n = self.nbf
        out = np.zeros(n ** 4)
        lib().oracle_eri_full(self._h, _p(out))
        return out.reshape(n, n, n, n)

    def shell_pairs(self):
        n = lib().oracle_shell_pairs(self._h, None, None)
        s1 = np.zeros(n, dtype=np.int32)
        s2 = np.zeros(n, dtype=np.int32)
        lib().oracle_shell_pairs(self._h, _p(s1), _p(s2))
        return s1, s2

    def schwarz(self):
        Q = np.zeros((self.nshell, self.nshell), order="F")
        lib().oracle_schwarz(self._h, _p(Q))
        return Q

    def contract(self, Xs, types, hers, thresh=1e-12, screen=True, nthreads=0):
        """twoBodyContract over a list of NB x NB matrices; returns AX list.

        Matrices are handled column-major as ChronusQ does; all of them must
        share one dtype (float64 or complex128)."""
        cplx = any(np.iscomplexobj(x) for x in Xs)
        dt = np.complex128 if cplx else np.float64
        Xf = [np.asfortranarray(x, dtype=dt) for x in Xs]
        AX = [np.zeros((self.nbf, self.nbf), dtype=dt, order="F") for _ in Xs]
        arr = (_Contraction * len(Xs))()
        for i, (x, a, t, h) in enumerate(zip(Xf, AX, types, hers)):
            arr[i].X, arr[i].AX = x.ctypes.data, a.ctypes.data
            arr[i].her, arr[i].type = int(h), int(t)
        rc = lib().oracle_two_body_contract(self._h, float(thresh), int(screen),
                                            len(Xs), C.cast(arr, C.c_void_p),
                                            int(cplx), int(nthreads))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        return AX

    def grad_contract(self, Xs, types, nthreads=0):
        """directScaffoldGrad (direct.hpp:2198-2580) over real Hermitian
        NB x NB densities: returns AX[g][m] for every nuclear coordinate
        g = 3 atom + xyz and matrix m (derivative-integral J / K)."""
        natom = len(self.Z)
        Xf = [np.asfortranarray(x, dtype=np.float64) for x in Xs]
        ngrad, nmat = 3 * natom, len(Xs)
        AX = [[np.zeros((self.nbf, self.nbf), order="F") for _ in Xs] for _ in range(ngrad)]
        arr = (_Contraction * (ngrad * nmat))()
        for g in range(ngrad):
            for m, (x, t) in enumerate(zip(Xf, types)):
                arr[g * nmat + m].X, arr[g * nmat + m].AX = x.ctypes.data, AX[g][m].ctypes.data
                arr[g * nmat + m].her, arr[g * nmat + m].type = 1, int(t)
        rc = lib().oracle_grad_contract(self._h, nmat, C.cast(arr, C.c_void_p), int(nthreads))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        return AX

    def prepare(self):
        lib().oracle_prepare(self._h)

    def stats(self):
        a, b = C.c_longlong(), C.c_longlong()
        lib().oracle_last_stats(self._h, C.byref(a), C.byref(b))
        return a.value, b.value
