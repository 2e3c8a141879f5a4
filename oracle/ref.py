"""ctypes access to MOPAC's own SCF routines compiled by build_ref.py —
TEST INFRASTRUCTURE, used only to check the C restatement (oracle.py).

available() is False when oracle/_ref/libmopac_ref.so was not built (the
reference sources are needed to build it)."""
import ctypes
import os

import numpy as np

from . import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_ref", "libmopac_ref.so")
V = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
_lib = None


def available():
    return os.path.exists(SO)


def load():
    global _lib
    if _lib is None:
        olib = oracle.load()
        olib.o_lapack_available()  # opens MKL with RTLD_GLOBAL: pulay's BLAS/LAPACK calls
        lib = ctypes.CDLL(SO, mode=os.RTLD_LAZY)
        lib.ref_setup.argtypes = [I, I, I, I]
        lib.ref_fock2.argtypes = [V, V, V, V, I, V, V, I]
        lib.ref_helect.argtypes = [I, V, V, V]
        lib.ref_helect.restype = D
        lib.ref_cnvg.argtypes = [V, V, V, I, V]
        lib.ref_diag.argtypes = [V, V, I, V, I]
        lib.ref_osinv.argtypes = [V, I, V]
        lib.ref_pulay_seq.argtypes = [I, I, V, V, V]
        _lib = lib
    return _lib


def P(a):
    return a.ctypes.data_as(V)


def setup(system, npulay=10):
    load().ref_setup(system.norbs, system.numat, len(system.w), npulay)


def fock2(system, f, ptot, p, mode=2):
    setup(system)
    f = np.array(f, dtype=np.float64, copy=True)
    load().ref_fock2(P(f), P(np.ascontiguousarray(ptot)), P(np.ascontiguousarray(p)),
                     P(np.ascontiguousarray(system.w)), system.numat,
                     P(np.ascontiguousarray(system.nfirst, dtype=np.int32)),
                     P(np.ascontiguousarray(system.nlast, dtype=np.int32)), mode)
    return f


def helect(n, p, h, f):
    return load().ref_helect(n, P(np.ascontiguousarray(p)), P(np.ascontiguousarray(h)),
                             P(np.ascontiguousarray(f)))


def cnvg(pnew, pold, p1, niter):
    pl = ctypes.c_double()
    load().ref_cnvg(P(pnew), P(pold), P(p1), niter, ctypes.byref(pl))
    return pl.value


def diag(fao, vector, nocc, eig):
    v = np.array(vector, dtype=np.float64, order="F", copy=True)
    load().ref_diag(P(np.ascontiguousarray(fao)), P(v), nocc, P(np.ascontiguousarray(eig)), v.shape[0])
    return v


def osinv(a):
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    d = ctypes.c_double()
    load().ref_osinv(P(a), a.shape[0], ctypes.byref(d))
    return a, d.value


def pulay_seq(n, f_seq, p_seq):
    f = np.array(f_seq, dtype=np.float64, copy=True)
    pl = np.zeros(f.shape[0])
    load().ref_pulay_seq(n, f.shape[0], P(f), P(np.ascontiguousarray(p_seq)), P(pl))
    return f, pl
