"""Python mirror of the reference's SCF routines over the HIP library.

Each method keeps the reference routine's name, argument meaning and packed
storage; all arithmetic runs in libmopachip.so on the GPU.
"""
import ctypes

import numpy as np

from . import lib as L


class Context:
    """One GPU context (mhip_init / mhip_finalize)."""

    def __init__(self, device=0):
        self.lib = L.load()
        self.h = ctypes.c_void_p()
        L.check(self.lib.mhip_init(device, ctypes.byref(self.h)))
        self.system = None

    def close(self):
        if self.h:
            L.check(self.lib.mhip_finalize(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def device_info(self):
        buf = ctypes.create_string_buffer(256)
        gib = ctypes.c_double()
        ncu = ctypes.c_int()
        L.check(self.lib.mhip_device_info(self.h, buf, 256, ctypes.byref(gib), ctypes.byref(ncu)))
        return buf.value.decode(), gib.value, ncu.value

    # -- system -----------------------------------------------------------
    def upload(self, system):
        s = system
        L.check(self.lib.mhip_upload_system(self.h, s.numat, L.ptr(s.nfirst), L.ptr(s.nlast), s.norbs,
                                            len(s.w), L.ptr(s.h), L.ptr(s.w)))
        self.system = system

    # -- single routines (reference names) --------------------------------
    def density(self, c, ndubl, occ=2.0):
        """density_for_GPU case 4/5: packed occ*C_occ C_occ^T."""
        n = self.system.norbs
        c = np.asfortranarray(c, dtype=np.float64)
        p = np.zeros(n * (n + 1) // 2)
        L.check(self.lib.mhip_density(self.h, L.ptr(c), ndubl, occ, L.ptr(p)))
        return p

    def fock2(self, f, ptot, p, mode=2):
        """fock2(f, ptot, p, w, w, w, numat, nfirst, nlast, mode): returns F + G."""
        f = np.array(f, dtype=np.float64, copy=True)
        L.check(self.lib.mhip_fock2(self.h, L.ptr(f), L.ptr(np.ascontiguousarray(ptot)),
                                    L.ptr(np.ascontiguousarray(p)), mode))
        return f

    def helect(self, p, h, f):
        e = ctypes.c_double()
        L.check(self.lib.mhip_helect(self.h, self.system.norbs, L.ptr(np.ascontiguousarray(p)),
                                     L.ptr(np.ascontiguousarray(h)), L.ptr(np.ascontiguousarray(f)),
                                     ctypes.byref(e)))
        return e.value

    def eigenvectors(self, fpacked):
        """eigenvectors_LAPACK: returns (C, eigs); fpacked is perturbed in place."""
        n = self.system.norbs
        c = np.zeros((n, n), order="F")
        e = np.zeros(n)
        L.check(self.lib.mhip_eigen(self.h, L.ptr(c), L.ptr(fpacked), L.ptr(e), n))
        return c, e

    def diag(self, fao, vector, nocc, eig):
        """Pseudo-diagonalization diag(fao, vector, nocc, eig, n, n); returns
        (rotated vector, number of rotations)."""
        n = self.system.norbs
        v = np.array(vector, dtype=np.float64, order="F", copy=True)
        nr = ctypes.c_longlong()
        L.check(self.lib.mhip_pseudo_diag(self.h, L.ptr(np.ascontiguousarray(fao)), L.ptr(v), nocc,
                                          L.ptr(np.ascontiguousarray(eig)), n, n, ctypes.byref(nr)))
        return v, nr.value

    # -- the whole loop ---------------------------------------------------
    def iter(self, pulay=False, scfcrt=1e-4, itrmax=2000, last=1, precision=0, pseudo_diag=False,
             adaptive_eigen=False, p_init=None, want=("p", "eig"), block_start=0, fragment=None,
             start_from_guess=False):
        s = self.system
        n = s.norbs
        o = L.ScfOpts()
        o.nclose = s.nclose
        o.pulay = int(bool(pulay))
        o.itrmax = itrmax
        o.scfcrt = scfcrt
        o.last = last
        o.precision = precision
        o.pseudo_diag = int(pseudo_diag)
        o.adaptive_eigen = int(bool(adaptive_eigen))
        o.enuclr = s.enuclr
        o.atheat = s.atheat
        pd = np.ascontiguousarray(s.pdiag)
        o.pdiag = pd.ctypes.data
        pin = None
        if p_init is not None:
            pin = np.ascontiguousarray(p_init, dtype=np.float64)
            o.p_init = pin.ctypes.data
        o.block_start = int(block_start)
        o.start_from_guess = int(bool(start_from_guess))
        frag = None
        if fragment is not None:
            frag = np.ascontiguousarray(fragment, dtype=np.int32)
            o.fragment = frag.ctypes.data
        r = L.ScfResult()
        out = {}
        p = np.zeros(s.mpack) if "p" in want else None
        f = np.zeros(s.mpack) if "f" in want else None
        c = np.zeros((n, n), order="F") if "c" in want else None
        e = np.zeros(n) if "eig" in want else None
        L.check(self.lib.mhip_scf(self.h, ctypes.byref(o), ctypes.byref(r), L.ptr(p), L.ptr(f),
                                  L.ptr(c), L.ptr(e)))
        out.update(escf=r.escf, ee=r.ee, niter=r.niter, converged=bool(r.converged),
                   n_full_diag=r.n_full_diag, n_pseudo_diag=r.n_pseudo_diag, seconds=r.seconds,
                   jacobi_sweeps=r.jacobi_sweeps, n_fragment_diag=r.n_fragment_diag,
                   n_split_iter=r.n_split_iter,
                   p=p, f=f, c=c, eig=e)
        return out
