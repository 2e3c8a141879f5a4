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


# ---- MOZYME sparse Fock build (numpy over the C fock2) -----------------------

A0, EV = 0.529177210903, 27.211386245988


def _unpack_sparse(v, nfirst, iorbs, nijbo, n):
    m = np.zeros((n, n))
    f0 = np.asarray(nfirst) - 1
    for a in range(len(iorbs)):
        for b in range(a + 1):
            o = nijbo[a, b]
            if o < 0:
                continue
            ra, rb = slice(f0[a], f0[a] + iorbs[a]), slice(f0[b], f0[b] + iorbs[b])
            if a == b:
                t = np.zeros((iorbs[a], iorbs[a]))
                t[np.tril_indices(iorbs[a])] = v[o:o + iorbs[a] * (iorbs[a] + 1) // 2]
                m[ra, ra] = t + t.T - np.diag(np.diag(t))
            else:
                blk = v[o:o + iorbs[a] * iorbs[b]].reshape(iorbs[a], iorbs[b])
                m[ra, rb] = blk
                m[rb, ra] = blk.T
    return m


def _pack_sparse(m, nfirst, iorbs, nijbo, size):
    out = np.zeros(size)
    f0 = np.asarray(nfirst) - 1
    for a in range(len(iorbs)):
        for b in range(a + 1):
            o = nijbo[a, b]
            if o < 0:
                continue
            ra, rb = slice(f0[a], f0[a] + iorbs[a]), slice(f0[b], f0[b] + iorbs[b])
            blk = m[ra, rb]
            if a == b:
                out[o:o + iorbs[a] * (iorbs[a] + 1) // 2] = blk[np.tril_indices(iorbs[a])]
            else:
                out[o:o + iorbs[a] * iorbs[b]] = blk.ravel()
    return out


def mozyme_fock2z(system, nijbo, mpack, h_sparse, am, ad, dd, p, feather=False, trunc1=7.0, trunc2=0.22):
    """buildf(f, partf, 0) with fock2z/fz2n for a discrete system, SEMIDIRECT
    (src/MOZYME/buildf.F90:16-52, src/MOZYME/fock2z.F90:753-1362).

    Stored pairs (nijbo >= 0): the same jab/kab/fock1 terms as the
    conventional fock2 (fock2z.F90:813-1073), so the C fock2 runs on the dense
    form of p with the integral blocks of the other pairs zeroed.  Other
    pairs: the point-charge monopole with the electron populations of
    chrge_for_MOZYME, and for nijbo = -2 the dipole terms, from am, ad, dd
    (fock2z.F90:1074-1200); with feather (l_feather, PM7) the monopole
    blends into ev a0 / r and the dipole terms fade by to_point's factor
    (src/integrals/mndod.F90:3462-3486).  `system` supplies nfirst, nlast,
    xyz and the hcore-order integrals of every pair."""
    n = system.norbs
    nf = np.asarray(system.nfirst)
    iorbs = (np.asarray(system.nlast) - nf + 1).astype(np.int64)
    lim = iorbs * (iorbs + 1) // 2
    w = np.array(system.w, dtype=np.float64)
    kr = 0
    for i in range(len(iorbs)):
        for j in range(i):
            m = lim[i] * lim[j]
            if nijbo[i, j] < 0:
                w[kr:kr + m] = 0.0
            kr += m
        kr += lim[i] * lim[i]

    class _S:
        pass
    s2 = _S()
    s2.norbs, s2.numat, s2.nfirst, s2.nlast, s2.w = n, system.numat, system.nfirst, system.nlast, w
    pk = _unpack_sparse(p, nf, iorbs, nijbo, n)[np.tril_indices(n)]
    g = fock2(s2, np.zeros(len(pk)), pk, 0.5 * pk)
    gd = np.zeros((n, n))
    gd[np.tril_indices(n)] = g
    gd = gd + gd.T - np.diag(np.diag(gd))
    f = np.asarray(h_sparse, dtype=np.float64) + _pack_sparse(gd, nf, iorbs, nijbo, mpack)
    xyz = np.asarray(system.xyz, dtype=np.float64)
    qe = np.array([sum(p[nijbo[a, a] + (k * (k + 1)) // 2 + k] for k in range(iorbs[a]))
                   for a in range(len(iorbs))])
    for A in range(len(iorbs)):
        mono = dg = 0.0
        sp = np.zeros(3)
        for B in range(len(iorbs)):
            if B == A or nijbo[A, B] >= 0:
                continue
            d = xyz[A] - xyz[B]
            r2 = d @ d
            aee = 0.5 / am[A] + 0.5 / am[B]
            w1 = EV / np.sqrt(r2 / A0 ** 2 + aee ** 2)
            r = np.sqrt(r2)
            cf = 1.0
            if feather:
                cf = 1.0 - np.exp(-(r - trunc1) ** 2 * trunc2) if r < trunc1 else 0.0
                w1 = w1 * cf + (1.0 - cf) * EV * A0 / r
            mono += qe[B] * w1
            if nijbo[A, B] != -2:
                continue
            u = d / r
            if abs(u[2]) > 0.99999999:
                u[2] = np.sign(u[2])
            if iorbs[A] > 1:
                ade = 0.5 / ad[A] + 0.5 / am[B]
                rp, rm = np.hypot(r / A0 + dd[A], ade), np.hypot(r / A0 - dd[A], ade)
                sp += qe[B] * (EV * (0.5 / rp - 0.5 / rm) * cf) * u
            if iorbs[B] > 1:
                ade = 0.5 / ad[B] + 0.5 / am[A]
                rp, rm = np.hypot(r / A0 + dd[B], ade), np.hypot(r / A0 - dd[B], ade)
                sb = EV * (0.5 / rp - 0.5 / rm) * cf
                j1 = nijbo[B, B]
                dg -= 2.0 * sb * (p[j1 + 1] * u[0] + p[j1 + 3] * u[1] + p[j1 + 6] * u[2])
        i1 = nijbo[A, A]
        for k in range(iorbs[A]):
            f[i1 + (k * (k + 1)) // 2 + k] += mono + dg
        if iorbs[A] > 1:
            f[i1 + 1] += sp[0]
            f[i1 + 3] += sp[1]
            f[i1 + 6] += sp[2]
    return f
