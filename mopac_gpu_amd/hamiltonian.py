"""PM7 Hamiltonian (H, W, core-core, isolated-atom reference) for H/C/N/O
molecules, from the library's host implementation of hcore."""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import lib as L

TORE = {1: 1, 6: 4, 7: 5, 8: 6}


@dataclass
class System:
    nat: np.ndarray
    xyz: np.ndarray
    norbs: int
    nfirst: np.ndarray  # 1-based, int32
    nlast: np.ndarray
    h: np.ndarray       # packed, mpack
    w: np.ndarray       # n2elec
    enuclr: float       # eV
    atheat: float       # kcal/mol (sum of atomic heats - isolated-atom energies)
    pdiag: np.ndarray
    nclose: int
    charge: int = 0
    e_disp: float = 0.0   # PM7 dispersion correction, kcal/mol
    e_hbond: float = 0.0  # PM7 hydrogen-bond correction, kcal/mol

    def hof(self, escf):
        """Heat of formation: the SCF's escf plus PM7's post-SCF corrections
        (src/corrections/post_scf_corrections.F90:112-114)."""
        return escf + self.e_disp + self.e_hbond

    @property
    def numat(self):
        return len(self.nat)

    @property
    def mpack(self):
        return self.norbs * (self.norbs + 1) // 2


def build(nat, xyz, charge=0):
    lib = L.load()
    nat = np.ascontiguousarray(nat, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    numat = len(nat)
    norbs = ctypes.c_int()
    n2 = ctypes.c_longlong()
    L.check(lib.mhip_pm7_sizes(numat, L.ptr(nat), ctypes.byref(norbs), ctypes.byref(n2)))
    n = norbs.value
    nf = np.zeros(numat, np.int32)
    nl = np.zeros(numat, np.int32)
    h = np.zeros(n * (n + 1) // 2)
    w = np.zeros(n2.value)
    pd = np.zeros(n)
    en = ctypes.c_double()
    at = ctypes.c_double()
    L.check(lib.mhip_pm7_hamiltonian(numat, L.ptr(nat), L.ptr(xyz), charge, L.ptr(nf), L.ptr(nl),
                                     L.ptr(h), L.ptr(w), ctypes.byref(en), ctypes.byref(at), L.ptr(pd)))
    nel = int(sum(TORE[int(z)] for z in nat)) - charge
    ed, eh = corrections(nat, xyz)
    return System(nat, xyz, n, nf, nl, h, w, en.value, at.value, pd, nel // 2, charge, ed, eh)


def corrections(nat, xyz):
    """PM7's dispersion and hydrogen-bond corrections (kcal/mol) of a
    geometry."""
    lib = L.load()
    nat = np.ascontiguousarray(nat, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    ed = ctypes.c_double()
    eh = ctypes.c_double()
    L.check(lib.mhip_pm7_corrections(len(nat), L.ptr(nat), L.ptr(xyz), ctypes.byref(ed), ctypes.byref(eh), None))
    return ed.value, eh.value
