"""MOZYME localized-MO (LMO) data in the reference's layout, and a builder
that makes it from a converged conventional SCF.

Layout (src/MOZYME/MOZYME_C.F90:116-162), kept exactly because the GPU
kernels and the reference routines read the same arrays:

  occupied LMO i:  atoms icocc[nncf[i] : nncf[i] + ncf[i]] (1-based atom
                   numbers), coefficients cocc[ncocc[i] : ...] atom after
                   atom (iorbs of each); nncf/ncocc are 0-based offsets and
                   each LMO owns the space up to the next LMO's offset (room
                   to grow in diagg2);
  virtual LMO i:   the same with nce, nnce, icvir, ncvir, cvir;
  sparse P, F:     atom-pair blocks at nijbo[a, b] (0-based, -1 when the pair
                   is not stored): a > b as an iorbs(a) x iorbs(b) block with
                   b's orbital fastest, a == b as the packed lower triangle.

The builder localizes the canonical occupied and virtual orbitals of a
converged SCF separately (Pipek-Mezey), drops atoms whose weight in an LMO is
below `keep`, and stores every atom pair within `rcut` (Angstrom).  It feeds
the parity tests of the MOZYME kernels with realistic LMOs; the reference's
own start (a Lewis structure, makvec) is not part of the SCF inner loop.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np


@dataclass
class LmoSet:
    numat: int
    norbs: int
    nfirst: np.ndarray      # 1-based
    nlast: np.ndarray       # 1-based
    nat: np.ndarray
    nijbo: np.ndarray       # (numat, numat) int32, symmetric, -1 = not stored
    mpack: int              # size of the sparse storage
    nocc: int
    nvir: int
    ncf: np.ndarray
    nncf: np.ndarray
    ncocc: np.ndarray
    icocc: np.ndarray
    cocc: np.ndarray
    nce: np.ndarray
    nnce: np.ndarray
    ncvir: np.ndarray
    icvir: np.ndarray
    cvir: np.ndarray
    eigs: np.ndarray        # occupied LMO energies
    eigv: np.ndarray        # virtual LMO energies
    fao: np.ndarray = field(default=None)    # sparse Fock matrix
    p: np.ndarray = field(default=None)      # sparse density matrix

    @property
    def iorbs(self):
        return (self.nlast - self.nfirst + 1).astype(np.int32)

    def copy(self):
        return LmoSet(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()})


# fillij.F90:96-126: MOZYME's default cutoffs for a discrete system (CUTOF2,
# CUTOF1, CUTOFS in Angstrom); with fewer than 30 atoms every pair is stored
NDDO_CUTOFF, DIPOLE_CUTOFF, OVERLAP_CUTOFF = 9.9, 10.0, 7.0


def build_nijbo(xyz, iorbs, rcut):
    """Offsets of the stored atom-pair blocks: pairs closer than rcut
    (Angstrom, r^2 < rcut^2 as fillij compares) and every atom with itself,
    laid out row by row (a, b <= a), a > b as iorbs[a] x iorbs[b], a == b as
    a packed triangle (fillij.F90:150-270, ijbo.F90:16-83)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    iorbs = np.asarray(iorbs, dtype=np.int64)
    numat = len(iorbs)
    d2 = ((xyz[:, None, :] - xyz[None, :, :]) ** 2).sum(axis=2)
    lower = np.tril(np.ones((numat, numat), dtype=bool))
    keep = lower & ((d2 < rcut * rcut) | np.eye(numat, dtype=bool))
    size = np.outer(iorbs, iorbs)
    np.fill_diagonal(size, iorbs * (iorbs + 1) // 2)
    a_idx, b_idx = np.nonzero(keep)          # row-major: a ascending, then b
    blk = size[a_idx, b_idx]
    off = np.concatenate(([0], np.cumsum(blk)[:-1]))
    nijbo = np.full((numat, numat), -1, dtype=np.int32)
    nijbo[a_idx, b_idx] = off
    nijbo[b_idx, a_idx] = off
    return nijbo, int(blk.sum())


def build_nijbo_shells(xyz, iorbs, r_near=None, r_dipole=None):
    """nijbo with the reference's three classes of pairs (fillij.F90:96-245):
    blocks stored for r^2 < r_near^2, -2 (point charge + dipole) for
    r^2 < r_dipole^2, -1 (point charge) beyond.  Left at None, the cutoffs are
    fillij's defaults: 9.9 and 10 Angstrom, and every pair stored when the
    system has fewer than 30 atoms; given, they act like CUTOF2= / CUTOF1=."""
    xyz = np.asarray(xyz, dtype=np.float64)
    small = len(xyz) < 30 and r_near is None and r_dipole is None
    r2_near = 1.0e6 if small else (NDDO_CUTOFF if r_near is None else r_near) ** 2
    r2_dip = 1.0e6 + 10.0 if small else (DIPOLE_CUTOFF if r_dipole is None else r_dipole) ** 2
    r2_dip = max(r2_dip, r2_near + 1.0e-4)  # fillij.F90:130-134
    nijbo, size = build_nijbo(xyz, iorbs, np.sqrt(r2_near))
    d2 = ((xyz[:, None, :] - xyz[None, :, :]) ** 2).sum(axis=2)
    nijbo[(nijbo < 0) & (d2 < r2_dip)] = -2
    return nijbo, size


def mozyme_integrals(system, nijbo):
    """The two-electron integrals as hcore_for_MOZYME stores them for
    SEMIDIRECT: the conventional hcore-order array with the blocks of pairs
    that are not stored dropped (per atom i: stored pairs j < i, then i's
    one-centre block)."""
    no = (np.asarray(system.nlast) - np.asarray(system.nfirst) + 1).astype(np.int64)
    lim = no * (no + 1) // 2
    w = np.asarray(system.w)
    keep = np.zeros(len(w), dtype=bool)
    kr = 0
    for i in range(len(no)):
        for j in range(i):
            n = lim[i] * lim[j]
            if nijbo[i, j] >= 0:
                keep[kr:kr + n] = True
            kr += n
        keep[kr:kr + lim[i] * lim[i]] = True
        kr += lim[i] * lim[i]
    assert kr == len(w)
    return w[keep]


def element_parameters(nat):
    """Per-atom (am, ad, dd) of PM7 (parameters_C after calpar)."""
    from . import lib as L
    lib = L.load()
    out = np.zeros(9)
    am, ad, dd = [], [], []
    for z in nat:
        L.check(0 if lib.mhip_pm7_element(int(z), L.ptr(out)) == 0 else 1)
        am.append(out[2])
        ad.append(out[3])
        dd.append(out[4])
    return np.array(am), np.array(ad), np.array(dd)


class MozymeSystem:
    """A system in MOZYME storage: sparse H, integrals for SEMIDIRECT, nijbo
    with near / dipole / point-charge pairs, per-atom parameters.  The
    cutoffs default to MOZYME's own (build_nijbo_shells; resonance blocks
    to overlap_cutoff = CUTOFS = 7 Angstrom)."""

    TRUNC1, TRUNC2 = 7.0, 0.22  # PM7's switch to point charges (molkst_C trunc_1, trunc_2)

    def __init__(self, system, r_near=None, r_dipole=None, nijbo=None, mpack=None, feather=True,
                 overlap_cutoff=OVERLAP_CUTOFF):
        self.s = system
        self.numat, self.norbs = system.numat, system.norbs
        self.nat = np.asarray(system.nat, dtype=np.int32)
        self.xyz = np.ascontiguousarray(system.xyz, dtype=np.float64)
        self.nfirst = np.asarray(system.nfirst, dtype=np.int32)
        self.nlast = np.asarray(system.nlast, dtype=np.int32)
        iorbs = self.nlast - self.nfirst + 1
        if nijbo is None:
            self.nijbo, self.mpack = build_nijbo_shells(self.xyz, iorbs, r_near, r_dipole)
        else:
            self.nijbo, self.mpack = np.asarray(nijbo, dtype=np.int32), int(mpack)
        # one-electron matrix and core-core energy as hcore_for_MOZYME forms them
        from . import lib as L
        self.h = np.zeros(self.mpack)
        en = ctypes.c_double()
        # diat's overlap limit under MOZYME is CUTOF1 (fillij.F90:96-101)
        if r_dipole is not None:
            cutof1 = r_dipole
        elif r_near is None and self.numat < 30:
            cutof1 = np.sqrt(1.0e6 + 10.0)
        else:
            cutof1 = DIPOLE_CUTOFF
        self.overlap_cutoff = float(overlap_cutoff)
        rc = L.load().mhip_pm7_hamiltonian_mozyme(
            self.numat, L.ptr(np.ascontiguousarray(self.nat)), L.ptr(self.xyz),
            L.ptr(np.asfortranarray(self.nijbo.astype(np.int32))), self.mpack, self.overlap_cutoff, float(cutof1),
            L.ptr(self.h), ctypes.byref(en))
        if rc != 0:
            raise ValueError("mhip_pm7_hamiltonian_mozyme: unsupported element or nijbo beyond mpack")
        self.enuclr = en.value
        # near-pair integrals in hcore order; the far pairs are point charges
        nijf = np.asfortranarray(self.nijbo.astype(np.int32))
        nw = ctypes.c_longlong()
        lib = L.load()
        L.check(lib.mhip_pm7_integrals_mozyme(self.numat, L.ptr(np.ascontiguousarray(self.nat)), L.ptr(self.xyz),
                                              L.ptr(nijf), ctypes.byref(nw), None))
        self.w = np.zeros(nw.value)
        L.check(lib.mhip_pm7_integrals_mozyme(self.numat, L.ptr(np.ascontiguousarray(self.nat)), L.ptr(self.xyz),
                                              L.ptr(nijf), ctypes.byref(nw), L.ptr(self.w)))
        self.am, self.ad, self.dd = element_parameters(self.nat)
        self.feather = bool(feather)

    @classmethod
    def from_geometry(cls, nat, xyz, r_near=None, r_dipole=None, charge=0, feather=True,
                      overlap_cutoff=OVERLAP_CUTOFF):
        """A MOZYME system straight from the geometry: the atom data and
        the sparse one- and two-electron terms only, without the
        conventional system's full W and packed H (which do not fit a
        protein-sized problem comfortably)."""
        from types import SimpleNamespace
        from . import lib as L
        from .hamiltonian import TORE
        nat = np.ascontiguousarray(nat, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        numat = len(nat)
        norbs = ctypes.c_int()
        n2 = ctypes.c_longlong()
        lib = L.load()
        L.check(lib.mhip_pm7_sizes(numat, L.ptr(nat), ctypes.byref(norbs), ctypes.byref(n2)))
        nf = np.zeros(numat, np.int32)
        nl = np.zeros(numat, np.int32)
        pd = np.zeros(norbs.value)
        at = ctypes.c_double()
        L.check(lib.mhip_pm7_atoms(numat, L.ptr(nat), charge, L.ptr(nf), L.ptr(nl), ctypes.byref(at), L.ptr(pd)))
        nel = int(sum(TORE[int(z)] for z in nat)) - charge
        system = SimpleNamespace(nat=nat, xyz=xyz, numat=numat, norbs=norbs.value, nfirst=nf, nlast=nl,
                                 atheat=at.value, pdiag=pd, nclose=nel // 2, charge=charge)
        return cls(system, r_near=r_near, r_dipole=r_dipole, feather=feather, overlap_cutoff=overlap_cutoff)

    def diagonal_density(self):
        """The standard starting density (pdiag on the diagonal) in sparse storage."""
        p = np.zeros(self.mpack)
        iorbs = self.nlast - self.nfirst + 1
        for a in range(self.numat):
            o = self.nijbo[a, a]
            for k in range(iorbs[a]):
                p[o + k * (k + 1) // 2 + k] = self.s.pdiag[self.nfirst[a] - 1 + k]
        return p

    def upload(self, ctx):
        from . import lib as L
        c = lambda a, t=np.float64: np.ascontiguousarray(a, dtype=t)  # noqa: E731
        self._keep = [c(self.nfirst, np.int32), c(self.nlast, np.int32),
                      np.asfortranarray(self.nijbo.astype(np.int32)), c(self.xyz), c(self.am), c(self.dd),
                      c(self.ad), c(self.h), c(self.w)]
        nf, nl, nij, xyz, am, dd, ad, h, w = self._keep
        sysc = L.MozymeSystemC(self.numat, self.norbs, self.mpack, nf.ctypes.data, nl.ctypes.data,
                               nij.ctypes.data, xyz.ctypes.data, am.ctypes.data, dd.ctypes.data, ad.ctypes.data,
                               h.ctypes.data, w.ctypes.data, len(w), int(self.feather), self.TRUNC1, self.TRUNC2)
        L.check(ctx.lib.mhip_mozyme_setup(ctx.h, ctypes.byref(sysc)))

    def scf(self, ctx, lmos, scfcrt=1e-2, itrmax=2000, thresh=1e-13, fmo_dim=0, print_level=0, heats_cap=0,
            order_free=False, fp32_fmo=False):
        """iter_for_MOZYME on the device from the LMOs of `lmos` (an LmoSet on
        this system's nijbo); returns the result and the updated LMOs."""
        from . import lib as L
        s = reference_capacity(lmos)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        a = dict(nfirst=i32(s.nfirst), nlast=i32(s.nlast), nijbo=np.asfortranarray(i32(s.nijbo)),
                 ncf=i32(s.ncf), nncf=i32(s.nncf), ncocc=i32(s.ncocc), icocc=i32(s.icocc),
                 cocc=np.ascontiguousarray(s.cocc, dtype=np.float64), nce=i32(s.nce), nnce=i32(s.nnce),
                 ncvir=i32(s.ncvir), icvir=i32(s.icvir), cvir=np.ascontiguousarray(s.cvir, dtype=np.float64))
        c = L.LmoSetC(numat=s.numat, norbs=s.norbs, nocc=s.nocc, nvir=s.nvir, mpack=s.mpack,
                      icocc_dim=len(a["icocc"]), cocc_dim=len(a["cocc"]), icvir_dim=len(a["icvir"]),
                      cvir_dim=len(a["cvir"]))
        for k, v in a.items():
            setattr(c, k, v.ctypes.data)
        heats = np.zeros(max(heats_cap, 1))
        o = L.MozymeOptsC(scfcrt=scfcrt, itrmax=itrmax, thresh=thresh, damp=1.0, fmo_dim=fmo_dim,
                          enuclr=self.enuclr, atheat=self.s.atheat, print_level=print_level,
                          heats=heats.ctypes.data if heats_cap > 0 else None, heats_cap=heats_cap,
                          order_free=int(order_free), fp32_fmo=int(fp32_fmo))
        r = L.MozymeResultC()
        p = np.zeros(self.mpack)
        f = np.zeros(self.mpack)
        L.check(ctx.lib.mhip_mozyme_scf(ctx.h, ctypes.byref(c), ctypes.byref(o), ctypes.byref(r), L.ptr(p),
                                        L.ptr(f)))
        out = s.copy()
        for k in ("ncf", "nncf", "ncocc", "icocc", "cocc", "nce", "nnce", "ncvir", "icvir", "cvir"):
            setattr(out, k, a[k].copy())
        out.p, out.fao = p, f
        return dict(escf=r.escf, ee=r.ee, niter=r.niter, converged=bool(r.converged), ovmax=r.ovmax,
                    seconds=r.seconds, p=p, f=f, lmos=out, heats=heats[:min(r.niter, heats_cap)].copy())

    def fock2z(self, ctx, p):
        from . import lib as L
        p = np.ascontiguousarray(p, dtype=np.float64)
        f = np.zeros(self.mpack)
        L.check(ctx.lib.mhip_mozyme_fock2z(ctx.h, L.ptr(p), L.ptr(f)))
        return f


def reference_capacity(s):
    """A copy of an LmoSet whose LMO arrays have the capacities the reference
    allocates (set_up_MOZYME_arrays.F90:53-58), the LMOs kept where they are
    and the extra room at the end (tidy spreads it over the LMOs)."""
    numat, norbs = s.numat, s.norbs
    ipad2 = min(numat + 10, int(round(numat ** 0.25 * 50)))
    ipad4 = int(round(ipad2 * norbs / numat))
    out = s.copy()

    def grow(a, n, dtype):
        b = np.zeros(max(n, len(a)), dtype=dtype)
        b[:len(a)] = a
        return b
    out.icocc = grow(s.icocc, ipad2 * s.nocc + 1000, np.int32)
    out.icvir = grow(s.icvir, ipad2 * s.nvir + 1000, np.int32)
    out.cocc = grow(s.cocc, ipad4 * s.nocc + 1000, np.float64)
    out.cvir = grow(s.cvir, ipad4 * s.nvir + 1000, np.float64)
    return out


def sparse_from_dense(m, nfirst, iorbs, nijbo, size):
    """Dense symmetric matrix (norbs x norbs) -> MOZYME sparse blocks."""
    out = np.zeros(size)
    numat = len(iorbs)
    f0 = nfirst - 1
    for a in range(numat):
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


def dense_from_sparse(v, nfirst, iorbs, nijbo, norbs):
    m = np.zeros((norbs, norbs))
    f0 = nfirst - 1
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


def pipek_mezey(c, nfirst, iorbs, sweeps=30, tol=1e-10):
    """Pipek-Mezey localization of the columns of c (orthogonal AO basis):
    maximize sum_i sum_A q_iA^2, q_iA = sum_{mu on A} c_mu,i^2, by 2x2
    rotations (Pipek & Mezey, J. Chem. Phys. 90, 4916 (1989))."""
    c = np.array(c, dtype=np.float64, copy=True)
    n = c.shape[1]
    f0 = nfirst - 1
    atom_of = np.repeat(np.arange(len(iorbs)), iorbs)
    nat = len(iorbs)

    def pops(x):
        return np.bincount(atom_of, weights=x, minlength=nat)

    for _ in range(sweeps):
        change = 0.0
        for i in range(n - 1):
            for j in range(i + 1, n):
                ci, cj = c[:, i], c[:, j]
                qii = pops(ci * ci)
                qjj = pops(cj * cj)
                qij = pops(ci * cj)
                a = np.sum(qij ** 2) - 0.25 * np.sum((qii - qjj) ** 2)
                b = np.sum(qij * (qii - qjj))
                if abs(a) < 1e-300 and abs(b) < 1e-300:
                    continue
                g = np.hypot(a, b)
                if a + g < 1e-14:
                    continue
                t = 0.25 * np.arctan2(b, -a)
                if abs(t) < 1e-12:
                    continue
                cs, sn = np.cos(t), np.sin(t)
                c[:, i], c[:, j] = cs * ci + sn * cj, -sn * ci + cs * cj
                change = max(change, abs(t))
        if change < tol:
            break
    return c


def _pack(cols, nfirst, iorbs, keep, pad_atoms):
    f0 = nfirst - 1
    nat = len(iorbs)
    ncf, nncf, ncoc, idx, coef = [], [], [], [], []
    ia = ic = 0
    for k in range(cols.shape[1]):
        x = cols[:, k]
        w = np.array([np.sum(x[f0[a]:f0[a] + iorbs[a]] ** 2) for a in range(nat)])
        atoms = [a for a in np.argsort(-w) if w[a] > keep]
        nncf.append(ia)
        ncoc.append(ic)
        ncf.append(len(atoms))
        for a in atoms:
            idx.append(a + 1)
            coef.extend(x[f0[a]:f0[a] + iorbs[a]])
        room = pad_atoms
        idx.extend([0] * room)
        coef.extend([0.0] * (4 * room))
        ia += len(atoms) + room
        ic += sum(iorbs[a] for a in atoms) + 4 * room
    return (np.array(ncf, np.int32), np.array(nncf, np.int32), np.array(ncoc, np.int32),
            np.array(idx, np.int32), np.array(coef))


def from_scf(system, c, f_dense, p_dense, nocc, keep=1e-4, rcut=8.0, pad_atoms=6, sweeps=30, r_dipole=None):
    """LMOs in MOZYME layout from canonical orbitals c (norbs x norbs); pairs
    within rcut are stored, with r_dipole the dipole shell is marked -2."""
    nfirst = np.asarray(system.nfirst, dtype=np.int32)
    nlast = np.asarray(system.nlast, dtype=np.int32)
    iorbs = nlast - nfirst + 1
    occ = pipek_mezey(c[:, :nocc], nfirst, iorbs, sweeps)
    vir = pipek_mezey(c[:, nocc:], nfirst, iorbs, sweeps)
    if r_dipole is None:
        nijbo, size = build_nijbo(np.asarray(system.xyz), iorbs, rcut)
    else:
        nijbo, size = build_nijbo_shells(np.asarray(system.xyz, dtype=np.float64), iorbs, rcut, r_dipole)
    o = _pack(occ, nfirst, iorbs, keep, pad_atoms)
    v = _pack(vir, nfirst, iorbs, keep, pad_atoms)
    s = LmoSet(numat=system.numat, norbs=system.norbs, nfirst=nfirst, nlast=nlast,
               nat=np.asarray(system.nat, dtype=np.int32), nijbo=nijbo, mpack=size, nocc=nocc,
               nvir=c.shape[1] - nocc, ncf=o[0], nncf=o[1], ncocc=o[2], icocc=o[3], cocc=o[4],
               nce=v[0], nnce=v[1], ncvir=v[2], icvir=v[3], cvir=v[4],
               eigs=np.zeros(nocc), eigv=np.zeros(c.shape[1] - nocc))
    s.fao = sparse_from_dense(f_dense, nfirst, iorbs, nijbo, size)
    s.p = sparse_from_dense(p_dense, nfirst, iorbs, nijbo, size)
    s.eigs = np.array([_energy(s, k, True, f_dense) for k in range(nocc)])
    s.eigv = np.array([_energy(s, k, False, f_dense) for k in range(s.nvir)])
    return s


VALENCE = {1: 1, 6: 4, 7: 5, 8: 6}


def _hybrids(u):
    """Orthonormal directed hybrids (s, px, py, pz) for m atoms with k bonds
    each, bond directions u (m x k x 3 unit vectors), closest in the
    least-squares sense to the sp3-like targets (1/2, sqrt(3)/2 u)
    (symmetric orthonormalization), and an orthonormal basis of the rest of
    the valence space: (m x 4 x k, m x 4 x (4 - k))."""
    u = np.asarray(u, dtype=np.float64)
    m, k = u.shape[:2]
    t = np.empty((m, 4, k))
    t[:, 0] = 0.5
    t[:, 1:] = 0.8660254037844386 * u.transpose(0, 2, 1)
    ev, vec = np.linalg.eigh(t.transpose(0, 2, 1) @ t)
    if m and ev.min() < 1e-6:
        raise ValueError("Lewis structure: linearly dependent bond directions")
    h = t @ ((vec * ev[:, None, :] ** -0.5) @ vec.transpose(0, 2, 1))
    proj = np.eye(4) - h @ h.transpose(0, 2, 1)
    # proj has eigenvalue 0 on the k hybrids and 1 on the rest (ascending)
    _, vec2 = np.linalg.eigh(proj)
    return h, vec2[:, :, k:]


def lewis_lmos(ms, fock=None, pad_atoms=6, scale=1.2):
    """Starting LMOs from a Lewis structure, the role makvec plays in the
    reference (src/MOZYME/makvec.F90:16-143): two-centre sigma bonds from
    directed hybrids, pi bonds between atoms left with an odd electron, and
    lone pairs, each bond with its antibond as the virtual partner.  In the
    orthonormal NDDO basis these are orthonormal by construction (every
    hybrid is used once).  Bonds are polarized by the 2 x 2 Fock matrix of
    their two hybrids when `fock` (sparse, on ms.nijbo; e.g. the Fock matrix
    of the diagonal starting density, as makvec builds it) is given, else
    equal weights.  Supports closed-shell H, C, N, O molecules whose
    bonding follows from the covalent radii."""
    from .geometry import COVALENT_RADIUS
    nat = np.asarray(ms.nat)
    xyz = np.asarray(ms.xyz, dtype=np.float64)
    numat = len(nat)
    iorbs = (ms.nlast - ms.nfirst + 1).astype(np.int64)
    rad = np.array([COVALENT_RADIUS[int(z)] for z in nat])
    from scipy.spatial import cKDTree
    pairs = cKDTree(xyz).query_pairs(2.0 * scale * rad.max(), output_type="ndarray")
    d = np.linalg.norm(xyz[pairs[:, 0]] - xyz[pairs[:, 1]], axis=1)
    pairs = pairs[d < scale * (rad[pairs[:, 0]] + rad[pairs[:, 1]])]
    nbr = [[] for _ in range(numat)]
    for a, b in pairs:
        nbr[a].append(int(b))
        nbr[b].append(int(a))
    # hybrids: hyb[a][b] = the orbital of atom a pointing at b; rest[a] = the others
    hyb, rest, elec = [dict() for _ in range(numat)], [None] * numat, np.zeros(numat, dtype=np.int64)
    heavy = {k: [] for k in range(1, 5)}
    for a in range(numat):
        k = len(nbr[a])
        if iorbs[a] == 1:
            if k != 1:
                raise ValueError(f"Lewis structure: hydrogen {a + 1} has {k} bonds")
            hyb[a][nbr[a][0]] = np.array([1.0])
            rest[a] = np.zeros((1, 0))
        else:
            if not 1 <= k <= 4:
                raise ValueError(f"Lewis structure: atom {a + 1} has {k} bonds")
            heavy[k].append(a)
        elec[a] = VALENCE[int(nat[a])] - k
    for k, atoms in heavy.items():
        if not atoms:
            continue
        # per-vector norms: the lone-pair basis spans a degenerate eigenspace,
        # so the directions are kept bit-stable against the one-atom form
        u = np.array([[(xyz[b] - xyz[a]) / np.linalg.norm(xyz[b] - xyz[a]) for b in nbr[a]]
                      for a in atoms])
        h, r = _hybrids(u)
        for ai, a in enumerate(atoms):
            for i, b in enumerate(nbr[a]):
                hyb[a][b] = h[ai, :, i]
            rest[a] = r[ai]

    tril = {n: np.tril_indices(n) for n in (1, 4)}
    diag_blk = {}

    def block(a, b):
        """F block rows on a, columns on b (dense, from the sparse storage)."""
        o = ms.nijbo[a, b]
        if a == b:
            t = diag_blk.get(a)
            if t is None:
                n = iorbs[a]
                t = np.zeros((n, n))
                t[tril[n]] = fock[o:o + n * (n + 1) // 2]
                t = t + t.T - np.diag(np.diag(t))
                diag_blk[a] = t
            return t
        if a > b:
            return fock[o:o + iorbs[a] * iorbs[b]].reshape(iorbs[a], iorbs[b])
        return block(b, a).T

    def polarize(reqs):
        """(c1, c2) of every bond: the lower eigenvector of the 2 x 2 Fock
        matrix of its two hybrids, [[ha F_aa ha, ha F_ab hb], [., hb F_bb hb]],
        in closed form (the rotation by t = atan2(2 fab, faa - fbb) / 2 puts
        the upper eigenvalue on (cos t, sin t), the lower on (-sin t, cos t)),
        signed so that c1 + c2 > 0 (c1 >= 0 when they cancel); equal weights
        without a Fock matrix.  Bonds with the same orbital counts are
        evaluated together."""
        m = len(reqs)
        if fock is None or m == 0:
            return np.full(m, np.sqrt(0.5)), np.full(m, np.sqrt(0.5))
        faa, fbb, fab = np.empty(m), np.empty(m), np.empty(m)
        groups = {}
        for k, (a, ha, b, hb) in enumerate(reqs):
            groups.setdefault((len(ha), len(hb)), []).append(k)
        for (na, nb), ks in groups.items():
            ha = np.array([reqs[k][1] for k in ks])
            hb = np.array([reqs[k][3] for k in ks])
            faa[ks] = np.einsum('ki,kij,kj->k', ha, np.array([block(reqs[k][0], reqs[k][0]) for k in ks]), ha)
            fbb[ks] = np.einsum('ki,kij,kj->k', hb, np.array([block(reqs[k][2], reqs[k][2]) for k in ks]), hb)
            fab[ks] = np.einsum('ki,kij,kj->k', ha, np.array([block(reqs[k][0], reqs[k][2]) for k in ks]), hb)
        t = 0.5 * np.arctan2(2.0 * fab, faa - fbb)
        u0, u1 = -np.sin(t), np.cos(t)
        flip = (u0 + u1 < 0) | ((u0 + u1 == 0) & (u0 < 0))
        return np.where(flip, -u0, u0), np.where(flip, -u1, u1)

    # the bonds (sigma here, pi below) as (a, hybrid of a, b, hybrid of b),
    # polarized together once all are known
    reqs = [(int(a), hyb[a][int(b)], int(b), hyb[b][int(a)]) for a, b in pairs]
    # pi bonds between bonded atoms that each keep an odd electron
    for a in range(numat):
        if elec[a] % 2 == 0:
            continue
        partners = [b for b in nbr[a] if elec[b] % 2 == 1 and rest[b].shape[1] > 0]
        if rest[a].shape[1] == 0 or not partners:
            raise ValueError(f"Lewis structure: no pi partner for atom {a + 1}")
        b = partners[0]
        ua, sv, vt = np.linalg.svd(rest[a][1:].T @ rest[b][1:])
        pa = rest[a] @ ua[:, 0]
        pb = rest[b] @ vt[0]
        rest[a] = rest[a] @ ua[:, 1:]
        rest[b] = rest[b] @ vt[1:].T
        elec[a] -= 1
        elec[b] -= 1
        reqs.append((a, pa, b, pb))
    occ, vir = [], []
    for (a, ha, b, hb), c1, c2 in zip(reqs, *polarize(reqs)):
        occ.append([(a, c1 * ha), (b, c2 * hb)])
        vir.append([(a, -c2 * ha), (b, c1 * hb)])
    for a in range(numat):
        n = rest[a].shape[1]
        if elec[a] != 2 * n:
            raise ValueError(f"Lewis structure: atom {a + 1} has {elec[a]} electrons for {n} lone pairs")
        for i in range(n):
            occ.append([(a, rest[a][:, i])])
    nocc = int(ms.s.nclose)
    if len(occ) != nocc or len(occ) + len(vir) != ms.norbs:
        raise ValueError(f"Lewis structure: {len(occ)} occupied / {len(vir)} virtual LMOs for "
                         f"{nocc} / {ms.norbs - nocc}")

    def pack(lmos):
        """Atoms of each LMO by descending weight (stable), each LMO followed
        by pad_atoms empty atom slots and 4 * pad_atoms zero coefficients."""
        ncf = np.array([len(parts) for parts in lmos], np.int32)
        idx, coef, ncc = [], [], np.empty(len(lmos), np.int64)
        ipad, cpad = [0] * pad_atoms, np.zeros(4 * pad_atoms)
        for k, parts in enumerate(lmos):
            parts = sorted(parts, key=lambda t: -float(t[1] @ t[1]))
            idx.extend(a + 1 for a, _ in parts)
            idx.extend(ipad)
            coef.extend(c for _, c in parts)
            coef.append(cpad)
            ncc[k] = sum(len(c) for _, c in parts) + 4 * pad_atoms
        nncf = np.concatenate(([0], np.cumsum(ncf + pad_atoms)[:-1])).astype(np.int32)
        ncoc = np.concatenate(([0], np.cumsum(ncc)[:-1])).astype(np.int32)
        return ncf, nncf, ncoc, np.array(idx, np.int32), np.concatenate(coef)

    o = pack(occ)
    v = pack(vir)
    return LmoSet(numat=numat, norbs=ms.norbs, nfirst=np.asarray(ms.nfirst, np.int32),
                  nlast=np.asarray(ms.nlast, np.int32), nat=nat.astype(np.int32), nijbo=ms.nijbo,
                  mpack=ms.mpack, nocc=nocc, nvir=len(vir), ncf=o[0], nncf=o[1], ncocc=o[2], icocc=o[3],
                  cocc=o[4], nce=v[0], nnce=v[1], ncvir=v[2], icvir=v[3], cvir=v[4], eigs=np.zeros(nocc),
                  eigv=np.zeros(len(vir)))


def lmo_dense(s, k, occupied):
    """Column k of the occupied (or virtual) LMOs as a dense norbs vector."""
    ncf, nncf, ncoc, idx, coef = ((s.ncf, s.nncf, s.ncocc, s.icocc, s.cocc) if occupied
                                  else (s.nce, s.nnce, s.ncvir, s.icvir, s.cvir))
    x = np.zeros(s.norbs)
    pos = ncoc[k]
    for a in idx[nncf[k]:nncf[k] + ncf[k]]:
        a0 = a - 1
        no = s.nlast[a0] - s.nfirst[a0] + 1
        x[s.nfirst[a0] - 1:s.nfirst[a0] - 1 + no] = coef[pos:pos + no]
        pos += no
    return x


def _energy(s, k, occupied, f_dense):
    x = lmo_dense(s, k, occupied)
    return float(x @ f_dense @ x)


class GpuMozyme:
    """The MOZYME kernels of libmopachip on an LmoSet, with the state the
    reference keeps between calls (mhip_diagg1_state / mhip_diagg2_state)."""

    def __init__(self, ctx, s, fmo_dim=None, thresh=1e-13, shift=0.0, damp=1.0, oldens=False, fp32=False):
        import ctypes
        from . import lib as L
        self.ctypes, self.L = ctypes, L
        self.ctx = ctx
        self.s = s.copy()
        self.fmo_dim = int(fmo_dim or max(1, s.nocc * s.nvir))
        self.nfmo = np.zeros(max(s.nvir, 1), dtype=np.int32)
        self.st1 = L.Diagg1State(reset=1, oldens=int(oldens), fp32_accumulate=int(fp32))
        self.st1.nfmo = self.nfmo.ctypes.data
        self.st2 = L.Diagg2State(reset=1, damp=damp, shift=shift, thresh=thresh)
        self.fmo = np.zeros(self.fmo_dim)
        self.ifmo = np.zeros((self.fmo_dim, 2), dtype=np.int32)
        self.nij = 0
        self._arrays()

    def _arrays(self):
        s = self.s
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        self.a = dict(nfirst=i32(s.nfirst), nlast=i32(s.nlast), nijbo=np.asfortranarray(i32(s.nijbo)),
                      ncf=i32(s.ncf), nncf=i32(s.nncf), ncocc=i32(s.ncocc), icocc=i32(s.icocc),
                      cocc=np.ascontiguousarray(s.cocc, dtype=np.float64), nce=i32(s.nce), nnce=i32(s.nnce),
                      ncvir=i32(s.ncvir), icvir=i32(s.icvir),
                      cvir=np.ascontiguousarray(s.cvir, dtype=np.float64))
        a = self.a
        self.c = self.L.LmoSetC(numat=s.numat, norbs=s.norbs, nocc=s.nocc, nvir=s.nvir, mpack=s.mpack,
                                icocc_dim=len(a["icocc"]), cocc_dim=len(a["cocc"]),
                                icvir_dim=len(a["icvir"]), cvir_dim=len(a["cvir"]))
        for k in ("nfirst", "nlast", "nijbo", "ncf", "nncf", "ncocc", "icocc", "cocc", "nce", "nnce", "ncvir",
                  "icvir", "cvir"):
            setattr(self.c, k, a[k].ctypes.data)
        self.eigs = np.array(s.eigs, dtype=np.float64)
        self.eigv = np.array(s.eigv, dtype=np.float64)

    def density(self, mode=0, partp=None):
        p = np.zeros(self.s.mpack)
        pp = None if partp is None else np.ascontiguousarray(partp, dtype=np.float64)
        self.L.check(self.L.load().mhip_mozyme_density(self.ctx.h, self.ctypes.byref(self.c), mode,
                                                       self.L.ptr(pp), self.L.ptr(p)))
        return p

    def diagg1(self, fao, p, idiagg):
        nij = self.ctypes.c_int()
        self.L.check(self.L.load().mhip_mozyme_diagg1(
            self.ctx.h, self.ctypes.byref(self.c), self.L.ptr(np.ascontiguousarray(fao)),
            self.L.ptr(np.ascontiguousarray(p)), self.L.ptr(self.eigs), self.L.ptr(self.eigv), idiagg,
            self.ctypes.byref(self.st1), self.fmo_dim, self.L.ptr(self.fmo), self.L.ptr(self.ifmo),
            self.ctypes.byref(nij)))
        self.nij = nij.value
        return dict(nij=self.nij, fmo=self.fmo[:self.nij].copy(), ifmo=self.ifmo[:self.nij].copy(),
                    tiny=self.st1.tiny, sumt=self.st1.sumt, ovmax=self.st1.ovmax, eigv=self.eigv.copy())

    def diagg2(self, idiagg):
        self.st2.tiny = self.st1.tiny
        self.L.check(self.L.load().mhip_mozyme_diagg2(
            self.ctx.h, self.ctypes.byref(self.c), self.L.ptr(self.eigs), self.L.ptr(self.eigv),
            self.L.ptr(self.fmo), self.L.ptr(self.ifmo), self.nij, idiagg, self.ctypes.byref(self.st2)))
        self.st1.tiny = self.st2.tiny
        return self.st2.sumb

    def tidy(self, occupied, thresh=1e-13):
        self.L.check(self.L.load().mhip_mozyme_tidy(self.ctx.h, self.ctypes.byref(self.c), int(occupied), thresh))

    def check(self, occupied):
        e = self.ctypes.c_double()
        self.L.check(self.L.load().mhip_mozyme_check(self.ctx.h, self.ctypes.byref(self.c), int(occupied),
                                                     self.ctypes.byref(e)))
        return e.value

    def get(self):
        s = self.s.copy()
        a = self.a
        s.ncf, s.icocc, s.cocc = a["ncf"].copy(), a["icocc"].copy(), a["cocc"].copy()
        s.nce, s.icvir, s.cvir = a["nce"].copy(), a["icvir"].copy(), a["cvir"].copy()
        s.nncf, s.ncocc, s.nnce, s.ncvir = a["nncf"].copy(), a["ncocc"].copy(), a["nnce"].copy(), a["ncvir"].copy()
        s.eigs, s.eigv = self.eigs.copy(), self.eigv.copy()
        return s
