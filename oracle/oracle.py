"""ctypes binding of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only the test suite, the smoke
check and the CPU-baseline leg of the benchmark import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_U64 = C.POINTER(C.c_uint64)
_I32 = C.POINTER(C.c_int32)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = C.CDLL(LIB)
        l.orc_create.restype = _P
        l.orc_create.argtypes = [C.c_int, C.c_double, _D, _D, C.c_int64, _U64, C.c_int64, _U64]
        l.orc_free.argtypes = [_P]
        l.orc_packed_size.restype = C.c_int64
        l.orc_packed_size.argtypes = [C.c_int]
        l.orc_pack_eri.argtypes = [C.c_int, _D, _D]
        l.orc_setup_from_fcidump.restype = C.c_int
        l.orc_setup_from_fcidump.argtypes = [C.c_int, C.c_int64, _D, _I32, _D, _D, _D]
        l.orc_table_nnz.restype = C.c_int64
        l.orc_table_nnz.argtypes = [_P, C.c_int]
        l.orc_table_get.argtypes = [_P, C.c_int, _I64, _I64]
        l.orc_hij_batch.argtypes = [_P, C.c_int64, _U64, _U64, _U64, _U64, _D]
        l.orc_hii.argtypes = [_P, _D]
        l.orc_sigma.argtypes = [_P, _D, _D, _D]
        l.orc_sigma_rows.argtypes = [_P, _D, _D, _D, C.c_int64, C.c_int64]
        l.orc_term_count.restype = C.c_int64
        l.orc_term_count.argtypes = [_P]
        l.orc_syev.argtypes = [C.c_int, _D, C.c_int, _D]
        l.orc_davidson.restype = C.c_int
        l.orc_davidson.argtypes = [_P, _D, _D, C.c_int, C.c_int, C.c_double, _D, _D, _D]
        l.orc_energy.restype = C.c_double
        l.orc_energy.argtypes = [_P, _D, _D]
        l.orc_rdm.argtypes = [_P, _D, _D, _D]
        l.orc_density.argtypes = [_P, _D, _D]
        l.orc_num_threads.restype = C.c_int
        _lib = l
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def pack_eri(eri):
    L = eri.shape[0]
    out = np.zeros(lib().orc_packed_size(L))
    lib().orc_pack_eri(L, _d(np.ascontiguousarray(eri, dtype=np.float64)), _d(out))
    return out


def setup_from_fcidump(L, values, idx):
    values = np.ascontiguousarray(values, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    I0 = np.zeros(1)
    h1 = np.zeros(L * L)
    packed = np.zeros(lib().orc_packed_size(L))
    rc = lib().orc_setup_from_fcidump(L, len(values), _d(values), idx.ctypes.data_as(_I32), _d(I0),
                                      _d(h1), _d(packed))
    if rc != 0:
        raise ValueError("bad FCIDUMP index")
    return float(I0[0]), h1.reshape(L, L), packed


class Problem:
    """Reference-convention problem: integrals + alpha/beta string sets + tables."""

    def __init__(self, norb, core, h1, astr, bstr, eri=None, packed=None):
        self.L = int(norb)
        self.h1 = np.ascontiguousarray(h1, dtype=np.float64)
        self.packed = pack_eri(eri) if packed is None else np.ascontiguousarray(packed, np.float64)
        self.a = np.ascontiguousarray(astr, dtype=np.uint64)
        self.b = np.ascontiguousarray(bstr, dtype=np.uint64)
        self.na, self.nb = len(self.a), len(self.b)
        self.core = float(core)
        self._p = lib().orc_create(self.L, self.core, _d(self.h1), _d(self.packed), self.na,
                                   self.a.ctypes.data_as(_U64), self.nb,
                                   self.b.ctypes.data_as(_U64))
        if not self._p:
            raise ValueError("orc_create failed")

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_free(self._p)
            self._p = None

    @property
    def ndet(self):
        return self.na * self.nb

    def table(self, which):
        """which: 0 Sa, 1 Da, 2 Sb, 3 Db -> (rowptr int64, cols int64)."""
        n = self.na if which < 2 else self.nb
        nnz = lib().orc_table_nnz(self._p, which)
        rp = np.zeros(n + 1, dtype=np.int64)
        cl = np.zeros(max(nnz, 1), dtype=np.int64)
        lib().orc_table_get(self._p, which, rp.ctypes.data_as(_I64), cl.ctypes.data_as(_I64))
        return rp, cl[:nnz]

    def hij(self, abra, bbra, aket, bket):
        abra, bbra, aket, bket = (np.ascontiguousarray(x, dtype=np.uint64) for x in (abra, bbra, aket, bket))
        out = np.zeros(len(abra))
        lib().orc_hij_batch(self._p, len(abra), abra.ctypes.data_as(_U64), bbra.ctypes.data_as(_U64),
                            aket.ctypes.data_as(_U64), bket.ctypes.data_as(_U64), _d(out))
        return out

    def hii(self):
        out = np.zeros(self.ndet)
        lib().orc_hii(self._p, _d(out))
        return out

    def sigma(self, c, hii=None):
        hii = self.hii() if hii is None else hii
        c = np.ascontiguousarray(c, dtype=np.float64)
        out = np.zeros(self.ndet)
        lib().orc_sigma(self._p, _d(hii), _d(c), _d(out))
        return out

    def sigma_rows(self, c, hii, ia0, ia1):
        c = np.ascontiguousarray(c, dtype=np.float64)
        out = np.zeros(self.ndet)
        lib().orc_sigma_rows(self._p, None if hii is None else _d(hii), _d(c), _d(out), ia0, ia1)
        return out

    def term_count(self):
        return lib().orc_term_count(self._p)

    def davidson(self, max_it=1, nb=10, eps=1e-4, w0=None, hii=None):
        hii = self.hii() if hii is None else hii
        W = np.zeros(self.ndet)
        if w0 is None:
            W[0] = 1.0
        else:
            W[:] = w0
        E0 = np.zeros(1)
        rn = np.zeros(1)
        energies = np.zeros(max_it * nb)
        nsig = lib().orc_davidson(self._p, _d(hii), _d(W), max_it, nb, eps, _d(E0), _d(rn), _d(energies))
        return dict(energy=float(E0[0]), resid=float(rn[0]), nsigma=nsig, W=W,
                    trajectory=energies[:nsig].copy())

    def energy(self, W, hii=None):
        hii = self.hii() if hii is None else hii
        return lib().orc_energy(self._p, _d(hii), _d(np.ascontiguousarray(W, np.float64)))

    def rdm(self, W):
        L = self.L
        one = np.zeros(2 * L * L)
        two = np.zeros(4 * L ** 4)
        lib().orc_rdm(self._p, _d(np.ascontiguousarray(W, np.float64)), _d(one), _d(two))
        return one.reshape(2, L * L), two.reshape(4, L ** 4)

    def density(self, W):
        out = np.zeros(2 * self.L)
        lib().orc_density(self._p, _d(np.ascontiguousarray(W, np.float64)), _d(out))
        return out


def syev(H):
    n = H.shape[0]
    U = np.asfortranarray(H.copy())
    E = np.zeros(n)
    lib().orc_syev(n, U.ctypes.data_as(_D), n, _d(E))
    return E, U
