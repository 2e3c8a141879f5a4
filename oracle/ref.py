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
        lib.mz_setup.argtypes = [I, I, I, V, V, V, V, D, D]
        lib.mz_set_lmos.argtypes = [I, I, V, V, V, I, V, I, V, V, V, V, I, V, I, V, I, V, V]
        lib.mz_get_lmos.argtypes = [I, I, V, V, V, V, V, V, V]
        lib.mz_density.argtypes = [V, I, I, V]
        lib.mz_diagg1.argtypes = [V, I, I, V, I, V, V, V, V, V, V]
        lib.mz_diagg2.argtypes = [I, I, V, I, I, V, V]
        lib.mz_helecz.argtypes = [V, V, V]
        lib.mz_helecz.restype = D
        lib.mz_cnvgz.argtypes = [V, V, V, V, V, I, V, I, V]
        lib.mz_fock2z.argtypes = [V, V, V, V, ctypes.c_longlong, V, V, V, D, D, I, D, D, V]
        lib.pm7_ref_init.argtypes = []
        lib.pm7_ref_element.argtypes = [I, V]
        lib.pm7_ref_pair.argtypes = [I, I, V, V, V, V, V, V, V]
        lib.pm7_ref_corrections.argtypes = [I, V, V, V]
        lib.pm7_ref_mozyme_hcore.argtypes = [I, V, V, D, D, D, V, I, V, V, V]
        lib.mz_tidy.argtypes = [I, V, V, I, V, I, V, V, I, D, V, V, V]
        lib.mz_check.argtypes = [I, V, V, V, I, V, V, I, V]
        lib.mz_eimp.argtypes = [V, V]
        lib.mz_iter_prologue.argtypes = [I, I, V, V, V]
        lib.mz_get_offsets.argtypes = [I, I, V, V, V, V]
        lib.mz_set_shift.argtypes = [D]
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


# ---- MOZYME (src/MOZYME/diagg1.F90, diagg2.F90, density_for_MOZYME.F90) ----

A0, EV = 0.529177210903, 27.211386245988
TRUNC1, TRUNC2 = 7.0, 0.22  # PM7's point-charge switch (the values the PM7 producer uses)


def mozyme_fock2z(ms, p, feather=False):
    """The reference's buildf(f, partf, 0) with fock2z (SEMIDIRECT storage) on
    a mopac_gpu_amd.mozyme.MozymeSystem and a sparse density p."""
    lib = load()
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    nij = np.asfortranarray(i32(ms.nijbo))
    lib.mz_setup(ms.numat, ms.norbs, ms.mpack, P(i32(ms.nfirst)), P(i32(ms.nlast)), P(i32(ms.nat)), P(nij),
                 1e-13, 0.0)
    f = np.zeros(ms.mpack)
    xyz = f64(ms.xyz)
    w = f64(ms.w)
    lib.mz_fock2z(P(xyz), P(f64(ms.am)), P(f64(ms.dd)), P(f64(ms.ad)), len(w), P(w), P(f64(p)), P(f64(ms.h)),
                  A0, EV, int(feather), TRUNC1, TRUNC2, P(f))
    return f

class Mozyme:
    """Loads an LmoSet (mopac_gpu_amd.mozyme) into the reference's module
    arrays and runs its MOZYME kernels; every call mutates the loaded set the
    way the reference does, and get() returns it."""

    def __init__(self, s, fmo_dim=None, thresh=1e-13, shift=0.0):
        lib = load()
        self.s = s.copy()
        self.fmo_dim = int(fmo_dim or max(1, s.nocc * s.nvir))
        self.i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        nij = np.asfortranarray(self.i32(s.nijbo))
        lib.mz_setup(s.numat, s.norbs, s.mpack, P(self.i32(s.nfirst)), P(self.i32(s.nlast)),
                     P(self.i32(s.nat)), P(nij), thresh, shift)
        self._keep = [self.i32(x) for x in (s.ncf, s.nncf, s.ncocc, s.icocc, s.nce, s.nnce, s.ncvir, s.icvir)]
        ncf, nncf, ncocc, icocc, nce, nnce, ncvir, icvir = self._keep
        self.cocc = np.ascontiguousarray(s.cocc, dtype=np.float64)
        self.cvir = np.ascontiguousarray(s.cvir, dtype=np.float64)
        eigs = np.ascontiguousarray(s.eigs, dtype=np.float64)
        p = np.ascontiguousarray(s.p, dtype=np.float64)
        lib.mz_set_lmos(s.nocc, s.nvir, P(ncf), P(nncf), P(ncocc), len(icocc), P(icocc), len(self.cocc),
                        P(self.cocc), P(nce), P(nnce), P(ncvir), len(icvir), P(icvir), len(self.cvir),
                        P(self.cvir), self.fmo_dim, P(eigs), P(p))
        self.eigv = np.array(s.eigv, dtype=np.float64)
        self.tiny = 0.0

    def density(self, mode=0, partp=None):
        s = self.s
        p = np.zeros(s.mpack)
        pp = np.zeros(s.mpack) if partp is None else np.ascontiguousarray(partp, dtype=np.float64)
        load().mz_density(P(p), mode, s.nocc, P(pp))
        return p

    def diagg1(self, fao, idiagg):
        s = self.s
        fmo = np.zeros(self.fmo_dim)
        ifmo = np.zeros((self.fmo_dim, 2), dtype=np.int32)
        nij = ctypes.c_int()
        tiny = ctypes.c_double(self.tiny)
        sumt, ovmax = ctypes.c_double(), ctypes.c_double()
        load().mz_diagg1(P(np.ascontiguousarray(fao)), s.nocc, s.nvir, P(self.eigv), idiagg, ctypes.byref(nij),
                         P(fmo), P(ifmo), ctypes.byref(tiny), ctypes.byref(sumt), ctypes.byref(ovmax))
        self.tiny = tiny.value
        k = nij.value
        return dict(nij=k, fmo=fmo[:k].copy(), ifmo=ifmo[:k].copy(), tiny=tiny.value, sumt=sumt.value,
                    ovmax=ovmax.value, eigv=self.eigv.copy())

    def diagg2(self, nij, idiagg):
        s = self.s
        tiny = ctypes.c_double(self.tiny)
        sumb = ctypes.c_double()
        load().mz_diagg2(s.nocc, s.nvir, P(self.eigv), nij, idiagg, ctypes.byref(tiny), ctypes.byref(sumb))
        self.tiny = tiny.value
        return sumb.value

    def helecz(self, p, h, f):
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        return load().mz_helecz(P(c(p)), P(c(h)), P(c(f)))

    def cnvgz(self, pnew, p, p1, p2, p3, niter, idiag, three_point=True):
        """In place on pnew, p, p1, p2, p3 (float64 arrays); returns pmax."""
        pm = ctypes.c_double()
        load().mz_cnvgz(P(pnew), P(p), P(p1), P(p2), P(p3), niter, P(self.i32(idiag)), int(three_point),
                        ctypes.byref(pm))
        return pm.value

    def prologue(self, p, f):
        """tidy, check and eimp as each iteration of iter_for_MOZYME starts;
        the module's p becomes eimp's array for the next diagg1."""
        s = self.s
        err = ctypes.c_int()
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        load().mz_iter_prologue(s.nocc, s.nvir, P(c(p)), P(c(f)), ctypes.byref(err))
        if err.value:
            raise RuntimeError("tidy: LMO storage too small")

    def get(self):
        s = self.s.copy()
        eigs = np.zeros(s.nocc)
        ncf = np.zeros(s.nocc, np.int32)
        nce = np.zeros(s.nvir, np.int32)
        icocc = np.zeros(len(s.icocc), np.int32)
        icvir = np.zeros(len(s.icvir), np.int32)
        cocc = np.zeros(len(s.cocc))
        cvir = np.zeros(len(s.cvir))
        load().mz_get_lmos(s.nocc, s.nvir, P(ncf), P(icocc), P(cocc), P(nce), P(icvir), P(cvir), P(eigs))
        s.ncf, s.nce, s.icocc, s.icvir, s.cocc, s.cvir, s.eigs = ncf, nce, icocc, icvir, cocc, cvir, eigs
        offs = [np.zeros(s.nocc, np.int32), np.zeros(s.nocc, np.int32), np.zeros(s.nvir, np.int32),
                np.zeros(s.nvir, np.int32)]
        load().mz_get_offsets(s.nocc, s.nvir, *[P(a) for a in offs])
        s.nncf, s.ncocc, s.nnce, s.ncvir = offs
        s.eigv = self.eigv.copy()
        return s


# ---- PM7 producer and post-SCF corrections (src/integrals/rotate.F90,
# h1elec.F90, ccrep.F90, src/models/calpar.F90, src/corrections/) ----------
# The MOZYME helpers above write per-element values into the same parameter
# module, so every PM7 call loads the PM7 set afresh.

NATORB = {1: 1, 6: 4, 7: 4, 8: 4}


ELEMENT_FIELDS = ("natorb", "tore", "am", "ad", "aq", "dd", "qq", "eisol", "eheat", "po1", "po2", "po3",
                  "po9", "ddp3", "uss", "upp", "gss", "gsp", "gpp", "gp2", "hsp")


def pm7_element(z):
    """Element z after calpar, as a dict of ELEMENT_FIELDS."""
    lib = load()
    lib.pm7_ref_init()
    out = np.zeros(21)
    lib.pm7_ref_element(int(z), P(out))
    return dict(zip(ELEMENT_FIELDS, out))


def pm7_pair(zi, zj, xi, xj, init=True):
    """rotate + h1elec for atom i (zi at xi) and atom j (zj at xj), Angstrom:
    (w (limij x limkl), e1b (limij), e2a (limkl), enuc eV, di (natorb_i x natorb_j))."""
    lib = load()
    if init:
        lib.pm7_ref_init()
    w = np.zeros(2025)
    e1b = np.zeros(45)
    e2a = np.zeros(45)
    en = ctypes.c_double()
    di = np.zeros((9, 9), order="F")
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    xj = np.ascontiguousarray(xj, dtype=np.float64)
    lib.pm7_ref_pair(int(zi), int(zj), P(xi), P(xj), P(w), P(e1b), P(e2a), ctypes.byref(en), P(di))
    ni, nj = NATORB[int(zi)], NATORB[int(zj)]
    li, lj = ni * (ni + 1) // 2, nj * (nj + 1) // 2
    return w[:li * lj].reshape(li, lj), e1b[:li].copy(), e2a[:lj].copy(), en.value, di[:ni, :nj].copy()


def pm7_hcore(nat, xyz):
    """H (packed), the two-centre W blocks and enuclr (eV) assembled from the
    reference's own pair routines in the order hcore visits pairs
    (src/integrals/hcore.F90:157-288): per atom i its pairs j < i; the
    resonance block into H(i, j), e1b onto atom i's diagonal block and e2a
    onto atom j's, Uss/Upp on the diagonal.  blocks[(i, j)] = w (limij x limkl)."""
    nat = [int(z) for z in nat]
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    el = {z: pm7_element(z) for z in set(nat)}
    no = [NATORB[z] for z in nat]
    nf = np.cumsum([0] + no[:-1])
    n = int(sum(no))
    h = np.zeros(n * (n + 1) // 2)
    pk = lambda r, c: r * (r + 1) // 2 + c  # noqa: E731  (r >= c, 0-based)
    for i, z in enumerate(nat):
        for a in range(no[i]):
            h[pk(nf[i] + a, nf[i] + a)] += el[z]["uss"] if a == 0 else el[z]["upp"]
    load().pm7_ref_init()
    blocks = {}
    enuc = 0.0
    for i in range(len(nat)):
        for j in range(i):
            w, e1b, e2a, en, di = pm7_pair(nat[i], nat[j], xyz[i], xyz[j], init=False)
            blocks[(i, j)] = w
            enuc += en
            for a in range(no[i]):
                for b in range(no[j]):
                    h[pk(nf[i] + a, nf[j] + b)] += di[a, b]
            for a in range(no[i]):
                for b in range(a + 1):
                    h[pk(nf[i] + a, nf[i] + b)] += e1b[a * (a + 1) // 2 + b]
            for a in range(no[j]):
                for b in range(a + 1):
                    h[pk(nf[j] + a, nf[j] + b)] += e2a[a * (a + 1) // 2 + b]
    return h, blocks, enuc, el


def pm7_corrections(nat, xyz):
    """(dispersion, hydrogen-bond energy, hydrogen bonds counted) in kcal/mol
    for a discrete molecule: PM6_DH_Dispersion and Hydrogen_bond_corrections
    as post_scf_corrections calls them for PM7."""
    lib = load()
    lib.pm7_ref_init()
    nat = np.ascontiguousarray(nat, dtype=np.int32)
    xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    out = np.zeros(3)
    lib.pm7_ref_corrections(len(nat), P(nat), P(xyz), P(out))
    return out


def pm7_mozyme_hcore(nat, xyz, cutof2=0.0, cutof1=0.0, cutofs=0.0):
    """fillij + hcore_for_MOZYME on a discrete molecule (SEMIDIRECT), the
    cutoffs MOZYME's defaults unless given (Angstrom, as the CUTOF2= /
    CUTOF1= / CUTOFS= keywords): (nijbo (numat x numat, 0-based offsets, -2,
    -1), mpack, h (sparse one-electron matrix), enuclr (eV))."""
    lib = load()
    lib.pm7_ref_init()
    nat = np.ascontiguousarray(nat, dtype=np.int32)
    xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    n = len(nat)
    nij = np.zeros((n, n), dtype=np.int32, order="F")
    mp = ctypes.c_int()
    en = ctypes.c_double()
    one = np.zeros(1)
    lib.pm7_ref_mozyme_hcore(n, P(nat), P(xyz), cutof2, cutof1, cutofs, P(nij), 0, ctypes.byref(mp), P(one),
                             ctypes.byref(en))
    h = np.zeros(mp.value)
    lib.pm7_ref_mozyme_hcore(n, P(nat), P(xyz), cutof2, cutof1, cutofs, P(nij), mp.value, ctypes.byref(mp), P(h),
                             ctypes.byref(en))
    return np.array(nij), mp.value, h, en.value


def _mz_setup_atoms(numat, norbs, mpack, nfirst, nlast, nat, nijbo):
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    keep = [i32(nfirst), i32(nlast), i32(nat), np.asfortranarray(i32(nijbo))]
    load().mz_setup(numat, norbs, mpack, *[P(a) for a in keep], 1e-13, 0.0)
    return keep


def mozyme_tidy(numat, norbs, iorbs, nc, nnc, ncmo, ic, c, thresh, mode=1):
    """tidy on one LMO set (copies in, copies out): returns (nc, nnc, ncmo,
    ic, c, ln, mn, err)."""
    iorbs = np.asarray(iorbs, dtype=np.int32)
    nl = np.cumsum(iorbs).astype(np.int32)
    keep = _mz_setup_atoms(numat, norbs, 1, nl - iorbs + 1, nl, np.ones(numat), np.zeros((numat, numat)))
    a = [np.array(x, dtype=np.int32) for x in (nc, ic, nnc, ncmo)]
    cc = np.array(c, dtype=np.float64)
    ln, mn, err = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    load().mz_tidy(len(a[0]), P(a[0]), P(a[1]), len(a[1]), P(cc), len(cc), P(a[2]), P(a[3]), mode, thresh,
                   ctypes.byref(ln), ctypes.byref(mn), ctypes.byref(err))
    del keep
    return a[0], a[2], a[3], a[1], cc, ln.value, mn.value, err.value


def mozyme_check(numat, norbs, iorbs, nc, nnc, ncmo, ic, c):
    """check on one LMO set: returns (c renormalized, ws = squared norms)."""
    iorbs = np.asarray(iorbs, dtype=np.int32)
    nl = np.cumsum(iorbs).astype(np.int32)
    keep = _mz_setup_atoms(numat, norbs, 1, nl - iorbs + 1, nl, np.ones(numat), np.zeros((numat, numat)))
    nv = len(nc)
    # check stops the job through mopend (not in this library) above 0.1
    err = 0.0
    for k in range(nv):
        n = int(sum(iorbs[a - 1] for a in ic[nnc[k]:nnc[k] + nc[k]]))
        err += abs(1.0 - float(np.sum(np.asarray(c[ncmo[k]:ncmo[k] + n]) ** 2)))
    if err > 0.1:
        raise ValueError(f"check: normalization error {err:.3g} > 0.1 would stop the reference")
    pad = lambda x: np.concatenate([np.asarray(x, dtype=np.int32), [0]]).astype(np.int32)  # noqa: E731
    a = [pad(nnc), pad(nc), np.array(ic, dtype=np.int32), pad(ncmo)]
    cc = np.array(c, dtype=np.float64)
    ws = np.zeros(nv)
    load().mz_check(nv, P(a[0]), P(a[1]), P(a[2]), len(a[2]), P(a[3]), P(cc), len(cc), P(ws))
    del keep
    return cc, ws


def mozyme_eimp(numat, norbs, nfirst, nlast, nijbo, mpack, f, p):
    """eimp on sparse F and P: returns the updated P."""
    keep = _mz_setup_atoms(numat, norbs, mpack, nfirst, nlast, np.ones(numat), nijbo)
    pp = np.array(p, dtype=np.float64)
    load().mz_eimp(P(np.ascontiguousarray(f, dtype=np.float64)), P(pp))
    del keep
    return pp


def mozyme_scf(ms, lmos, scfcrt=1e-2, itrmax=2000, fmo_dim=None, on_iteration=None):
    """iter_for_MOZYME's single-point loop driven over the reference's own
    routines (tidy, check, eimp, diagg1, diagg2, density_for_MOZYME, cnvgz,
    fz2n, helecz), in the order of src/MOZYME/iter_for_MOZYME.F90:220-640,
    with isitsc's test (isitsc.F90:31-46, emin = 0) and the loop's
    three-point extrapolation and level-shift schedule (:617-630), from the
    LmoSet lmos (taken with the reference's capacities) on the MozymeSystem
    ms.  Returns dict(escf, niter, converged, heats, ovmax)."""
    from mopac_gpu_amd import mozyme
    lib = load()
    L = mozyme.reference_capacity(lmos)
    L.eigs, L.eigv = np.zeros(L.nocc), np.zeros(L.nvir)
    L.p, L.fao = np.zeros(L.mpack), np.zeros(L.mpack)
    if fmo_dim is None:
        fmo_dim = min(L.nocc * L.nvir, 50 * max(L.nocc, L.nvir) * 64)
    R = Mozyme(L, fmo_dim=fmo_dim)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    xyz, am, dd, ad, w, h = f64(ms.xyz), f64(ms.am), f64(ms.dd), f64(ms.ad), f64(ms.w), f64(ms.h)
    fpc9 = 23.060547830619029  # eV -> kcal/mol (fpc(9) of the 2018 CODATA set)

    def fock(p):
        f = np.zeros(ms.mpack)
        lib.mz_fock2z(P(xyz), P(am), P(dd), P(ad), len(w), P(w), P(f64(p)), P(h), A0, EV, 1, TRUNC1, TRUNC2,
                      P(f))
        return f

    def heat(p, f):
        return (R.helecz(p, h, f) + ms.enuclr) * fpc9 + ms.s.atheat

    iorbs = (L.nlast - L.nfirst + 1).astype(np.int64)
    idiag = np.array([L.nijbo[a, a] + k * (k + 1) // 2 + k + 1 for a in range(L.numat) for k in range(iorbs[a])],
                     dtype=np.int32)
    p = R.density(0)
    f = fock(p)
    escf = heat(p, f)
    pold = np.zeros(L.mpack)
    p1, p2, p3 = np.zeros(L.norbs), np.zeros(L.norbs), np.zeros(L.norbs)
    eold, shift, niter, idiagg, lstart = 0.0, 0.0, 0, 0, 0
    three, scf1, okscf = True, False, False
    heats, ovmaxes = [], []
    while True:
        R.prologue(p, f)
        r = R.diagg1(f, idiagg)
        lib.mz_set_shift(shift)
        R.diagg2(r["nij"], idiagg)
        p = R.density(0)
        idiagg += 1
        niter += 1
        if three:
            R.cnvgz(p, pold, p1, p2, p3, niter, idiag, True)
        f = fock(p)
        escf = heat(p, f)
        ediff = escf - eold
        eold = escf
        ovmax = r["ovmax"]
        heats.append(escf)
        ovmaxes.append(ovmax)
        if on_iteration:
            on_iteration(niter, escf, ediff, ovmax)
        ok_now = ovmax < 5.0 * scfcrt and abs(ediff) < scfcrt
        if (ok_now and scf1) or niter > itrmax:
            okscf = True
        else:
            scf1 = ok_now
        if okscf and niter > 3:
            break
        if three:
            if niter % 3 == 2 and abs(ediff) < 0.1:
                three, lstart, shift = False, niter, 0.0
        elif ediff > 0.0 and shift < 11.0 and niter > lstart + 2:
            shift += 2.0
            lstart = niter
    return dict(escf=escf, niter=niter, converged=niter <= itrmax, heats=np.array(heats),
                ovmax=np.array(ovmaxes))
