"""ctypes access to the CPU FP64 restatement (oracle_scf.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
bench's cpu_baseline leg, as the checker — never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")
_lib = None
V = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return SO


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        lib = ctypes.CDLL(SO)
        lib.o_iter_flat.argtypes = [I, I, I, V, V, V, V, V, V, D, D, I, I, D, I, V, V, V, V, V, V, I, V, V, V]
        lib.o_iter_flat.restype = I
        lib.o_fock2.argtypes = [I, I, V, V, V, V, V, V]
        lib.o_fock2_general.argtypes = [I, V, V, V, V, V, V]
        lib.o_helect.argtypes = [I, V, V, V]
        lib.o_helect.restype = D
        lib.o_density.argtypes = [I, I, D, V, V]
        lib.o_eigenvectors.argtypes = [I, V, V, V]
        lib.o_eigenvectors.restype = I
        lib.o_cnvg.argtypes = [I, V, V, V, I, V]
        lib.o_diag.argtypes = [V, V, I, V, I]
        lib.o_diag.restype = ctypes.c_longlong
        lib.o_osinv.argtypes = [V, I, V]
        lib.o_pulay_seq.argtypes = [I, I, I, V, V, V]
        lib.o_lapack_available.restype = I
        lib.o_set_threads.argtypes = [I]
        _lib = lib
    return _lib


def P(a):
    return a.ctypes.data_as(V) if a is not None else None


def fock2(system, f, ptot, p, general=False):
    lib = load()
    f = np.array(f, dtype=np.float64, copy=True)
    if general:
        lib.o_fock2_general(system.numat, P(system.nfirst), P(system.nlast), P(f), P(ptot), P(p), P(system.w))
    else:
        lib.o_fock2(system.norbs, system.numat, P(system.nfirst), P(system.nlast), P(f), P(ptot), P(p),
                    P(system.w))
    return f


def helect(n, p, h, f):
    return load().o_helect(n, P(np.ascontiguousarray(p)), P(np.ascontiguousarray(h)), P(np.ascontiguousarray(f)))


def density(c, ndubl, occ=2.0):
    n = c.shape[0]
    c = np.asfortranarray(c)
    p = np.zeros(n * (n + 1) // 2)
    load().o_density(n, ndubl, occ, P(c), P(p))
    return p


def eigenvectors(fpacked):
    """Returns (C, eig); fpacked gets the reference's in-place i*1e-10 shift."""
    n = int((np.sqrt(8 * len(fpacked) + 1) - 1) // 2)
    c = np.zeros((n, n), order="F")
    e = np.zeros(n)
    load().o_eigenvectors(n, P(c), P(fpacked), P(e))
    return c, e


def cnvg(pnew, pold, p1, niter):
    pl = ctypes.c_double()
    load().o_cnvg(len(p1), P(pnew), P(pold), P(p1), niter, ctypes.byref(pl))
    return pl.value


def diag(fao, vector, nocc, eig):
    v = np.array(vector, dtype=np.float64, order="F", copy=True)
    n = v.shape[0]
    nr = load().o_diag(P(np.ascontiguousarray(fao)), P(v), nocc, P(np.ascontiguousarray(eig)), n)
    return v, nr


def osinv(a):
    a = np.array(a, dtype=np.float64, order="F", copy=True)
    d = ctypes.c_double()
    load().o_osinv(P(a), a.shape[0], ctypes.byref(d))
    return a, d.value


def pulay_seq(n, f_seq, p_seq, npulay=10):
    """len(f_seq) Pulay steps from a fresh start; returns (extrapolated F's, PLs)."""
    f = np.array(f_seq, dtype=np.float64, copy=True)
    pl = np.zeros(f.shape[0])
    load().o_pulay_seq(n, npulay, f.shape[0], P(f), P(np.ascontiguousarray(p_seq)), P(pl))
    return f, pl


def iterate(system, pulay=False, scfcrt=1e-4, itrmax=2000, last=1, p_init=None, trace_cap=2000):
    lib = load()
    s = system
    n = s.norbs
    p = np.zeros(s.mpack)
    f = np.zeros(s.mpack)
    c = np.zeros((n, n), order="F")
    e = np.zeros(n)
    tr = np.zeros(trace_cap)
    trp = np.zeros(trace_cap)
    escf = ctypes.c_double()
    ee = ctypes.c_double()
    nit = ctypes.c_int()
    pin = None if p_init is None else np.ascontiguousarray(p_init, dtype=np.float64)
    rc = lib.o_iter_flat(n, s.numat, s.nclose, P(s.nfirst), P(s.nlast), P(s.h), P(s.w), P(s.pdiag),
                         P(pin), s.enuclr, s.atheat, int(pulay), itrmax, scfcrt, last, P(p), P(f), P(c),
                         P(e), P(tr), P(trp), trace_cap, ctypes.byref(escf), ctypes.byref(ee),
                         ctypes.byref(nit))
    k = nit.value
    return dict(rc=rc, converged=(rc == 0), escf=escf.value, ee=ee.value, niter=k, p=p, f=f, c=c, eig=e,
                trace=tr[:min(k, trace_cap)], trace_pl=trp[:min(k, trace_cap)])
