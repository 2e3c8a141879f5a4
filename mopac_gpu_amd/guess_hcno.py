"""Learned starting density for molecules of H, C, N and O.

The pair-block network of the water guess (guess.py), generalized to any
element pair: an MLP (3 x 256 hidden units, BF16 weights, FP32 accumulation;
the same device kernels, k_mlp.hip) maps an atom pair -- its element-pair
type, the pair geometry in the first atom's frame and both atoms' local
environments within 4 A -- to that pair's density-matrix block.

Atom frames.  A heavy atom's frame is built from its bonded neighbours
(bonded: closer than 1.2 x the sum of covalent radii), taken heaviest first,
then nearest: e1 points to the first, e2 is the second's direction made
orthogonal to e1, e3 = e1 x e2.  An atom with one bonded neighbour (a
carbonyl O) takes e2 from that neighbour's own neighbours, so its frame lies
in the plane of the group it ends; with no usable neighbour the nearest
atoms within 4 A stand in, then a fixed axis.  A hydrogen takes the frame of
the heavy atom it is bonded to (its own block is an s orbital, so the frame
only orients the pair geometry).

Blocks.  The block P_AB (up to 4 x 4: s, px, py, pz) is predicted in the
frames of A and B, L_AB = U_A^T P_AB U_B with U = diag(1, R^T), so it turns
with the molecule.  Pairs are the atom itself and every atom within 4 A;
blocks beyond are zero.  The assembled density is renormalized per covalent
fragment so that its diagonal holds the fragment's valence electrons.  The
SCF starts from it as the reference starts from an OLDENS density
(iter.F90:151-159); atoms other than H, C, N, O, or no trained model, leave
the standard diagonal guess (moldat.F90:731-790) in place.

Long range.  Beyond 4 A an atom sees the rest of the molecule through the
potential and field (in its frame) of point charges: every atom carries the
mean Mulliken charge of its bonding type (element and bonded-neighbour
counts per element) over the training densities, a table stored with the
network.  This carries the polarization of a long helix's middle by the
aligned peptide dipoles, which the 4 A descriptors cannot see.

Descriptors and assembly run on the host (numpy, vectorized over pairs);
the network runs on the device.
"""
import os

import numpy as np
from scipy.spatial import cKDTree

from . import geometry
from . import guess as G

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "hcno_guess.npz")

ELEM = {1: 0, 6: 1, 7: 2, 8: 3}
VALENCE = {1: 1.0, 6: 4.0, 7: 5.0, 8: 6.0}
NIN = 168
NOUT = 16
NTYPES = 20  # self H, C, N, O; then ordered element pairs 4 + 4 e(A) + e(B)
RCUT = 4.0
MUS = np.linspace(1.0, 4.0, 5)
ETA = 3.0
BOND_SCALE = 1.2
NATOM = 56  # per-atom descriptor length


def supported(nat):
    return all(int(z) in ELEM for z in np.asarray(nat).ravel())


def layer_dims():
    return [NIN] + [G.NHID] * G.NLAYERS + [NOUT]


def bonded_neighbours(nat, xyz):
    """Per atom: bonded neighbours, heaviest first, then nearest (distances
    to 0.01 A), then the one lying further toward the centroid of the atoms
    within 4 A (projections to 0.001 A), then the lower index: an order that
    turns with the molecule, also for the equal bonds of a symmetric group."""
    nat = np.asarray(nat)
    rad = np.array([geometry.COVALENT_RADIUS[int(z)] for z in nat])
    tree = cKDTree(xyz)
    near = tree.query_ball_point(xyz, RCUT)
    cen = np.array([xyz[lst].mean(axis=0) - xyz[i] for i, lst in enumerate(near)])
    pairs = tree.query_pairs(2.0 * BOND_SCALE * rad.max(), output_type="ndarray")
    nb = [[] for _ in range(len(nat))]
    if len(pairs):
        d = np.linalg.norm(xyz[pairs[:, 0]] - xyz[pairs[:, 1]], axis=1)
        keep = d < BOND_SCALE * (rad[pairs[:, 0]] + rad[pairs[:, 1]])
        for (i, j), r in zip(pairs[keep], d[keep]):
            u = (xyz[j] - xyz[i]) / r
            nb[i].append((-int(nat[j]), round(float(r), 2), round(-float(np.dot(u, cen[i])), 3), j))
            nb[j].append((-int(nat[i]), round(float(r), 2), round(float(np.dot(u, cen[j])), 3), i))
    return [[x[-1] for x in sorted(v)] for v in nb], tree


def _frame(v1, v2):
    e1 = v1 / np.linalg.norm(v1)
    w = v2 - np.dot(v2, e1) * e1
    e2 = w / np.linalg.norm(w)
    return np.stack([e1, e2, np.cross(e1, e2)])


def atom_frames(nat, xyz, nbrs, tree):
    """R[i] (3 x 3, rows e1, e2, e3) for every atom."""
    nat = np.asarray(nat)
    n = len(nat)
    R = np.tile(np.eye(3), (n, 1, 1))
    heavy = nat != 1
    for i in np.flatnonzero(heavy):
        cands = list(nbrs[i])
        if len(cands) == 1:
            cands += [j for j in nbrs[cands[0]] if j != i]
        if len(cands) < 2:
            near = tree.query_ball_point(xyz[i], RCUT)
            near = sorted((np.linalg.norm(xyz[j] - xyz[i]), j) for j in near if j != i and j not in cands)
            cands += [j for _, j in near]
        v = [xyz[j] - xyz[i] for j in cands]
        if not v:
            continue
        v1 = v[0]
        v2 = None
        for w in v[1:]:
            if np.linalg.norm(np.cross(v1, w)) > 0.1 * np.linalg.norm(v1) * np.linalg.norm(w):
                v2 = w
                break
        if v2 is None:
            axis = np.eye(3)[int(np.argmin(np.abs(v1)))]
            v2 = axis
        R[i] = _frame(v1, v2)
    for i in np.flatnonzero(~heavy):
        hv = [j for j in nbrs[i] if nat[j] != 1]
        if hv:
            R[i] = R[hv[0]]
    return R


def atom_descriptors(nat, xyz, nbrs, tree, R):
    """Per atom (56): element one-hot, bonded counts per element, radial
    environment per element (5 Gaussians, cosine cutoff at 4 A), the first
    four bonded neighbours' positions in the atom's frame (/4 A) with their
    atomic numbers (/8), and the same for up to three further neighbours of
    the first one (the next shell along the bond the frame starts on)."""
    nat = np.asarray(nat)
    n = len(nat)
    e = np.array([ELEM[int(z)] for z in nat])
    D = np.zeros((n, NATOM))
    D[np.arange(n), e] = 1.0
    for i in range(n):
        for j in nbrs[i]:
            D[i, 4 + e[j]] += 1.0
    pairs = tree.query_pairs(RCUT, output_type="ndarray")
    if len(pairs):
        i, j = pairs[:, 0], pairs[:, 1]
        r = np.linalg.norm(xyz[i] - xyz[j], axis=1)
        g = np.exp(-ETA * (r[:, None] - MUS[None, :]) ** 2) * (0.5 * (np.cos(np.pi * r / RCUT) + 1.0))[:, None]
        for k in range(5):
            np.add.at(D, (i, 8 + 5 * e[j] + k), g[:, k])
            np.add.at(D, (j, 8 + 5 * e[i] + k), g[:, k])
    for i in range(n):
        for s, j in enumerate(nbrs[i][:4]):
            D[i, 28 + 3 * s: 31 + 3 * s] = R[i] @ (xyz[j] - xyz[i]) / RCUT
            D[i, 40 + s] = nat[j] / 8.0
        if nbrs[i]:
            nxt = [j for j in nbrs[nbrs[i][0]] if j != i][:3]
            for s, j in enumerate(nxt):
                D[i, 44 + 3 * s: 47 + 3 * s] = R[i] @ (xyz[j] - xyz[i]) / RCUT
                D[i, 53 + s] = nat[j] / 8.0
    return D


def atom_types(nat, nbrs):
    """Bonding type code: element, then bonded-neighbour counts per element."""
    e = np.array([ELEM[int(z)] for z in nat])
    cnt = np.zeros((len(nat), 4), dtype=np.int64)
    for i, lst in enumerate(nbrs):
        for j in lst:
            cnt[i, e[j]] += 1
    cnt = np.minimum(cnt, 4)
    return e + 4 * (cnt[:, 0] + 5 * (cnt[:, 1] + 5 * (cnt[:, 2] + 5 * cnt[:, 3])))


def charges_from_density(system, p):
    """Mulliken charges: valence electrons minus diagonal populations."""
    nf = (system.nfirst - 1).astype(np.int64)
    no = (system.nlast - system.nfirst + 1).astype(np.int64)
    orbs = np.arange(system.norbs)
    d = p[orbs * (orbs + 1) // 2 + orbs]
    pop = np.add.reduceat(d, nf) if len(nf) else np.zeros(0)
    pop = np.where(no > 0, pop, 0.0)
    return np.array([VALENCE[int(z)] for z in system.nat]) - pop


def field_descriptors(xyz, R, q):
    """Per atom: potential and field (e / A, e / A^2; field in the atom's
    frame) of the charges q of every atom farther than 4 A."""
    n = len(xyz)
    out = np.zeros((n, 4))
    for i0 in range(0, n, 512):
        i1 = min(n, i0 + 512)
        d = xyz[i0:i1, None, :] - xyz[None, :, :]
        r = np.linalg.norm(d, axis=2)
        w = np.where(r > RCUT, q[None, :] / np.maximum(r, RCUT), 0.0)
        out[i0:i1, 0] = w.sum(axis=1)
        e = np.einsum("ij,ijk->ik", w / np.maximum(r, RCUT) ** 2, d)
        out[i0:i1, 1:] = np.einsum("ikl,il->ik", R[i0:i1], e)
    return out


def pair_list(nat, tree):
    """Every (A, B) with A >= B: the atom itself and every atom within 4 A."""
    n = len(nat)
    pairs = tree.query_pairs(RCUT, output_type="ndarray")
    hi = np.maximum(pairs[:, 0], pairs[:, 1]) if len(pairs) else np.zeros(0, np.int64)
    lo = np.minimum(pairs[:, 0], pairs[:, 1]) if len(pairs) else np.zeros(0, np.int64)
    o = np.lexsort((lo, hi))
    a = np.concatenate([np.arange(n), hi[o]])
    b = np.concatenate([np.arange(n), lo[o]])
    return a.astype(np.int64), b.astype(np.int64)


def prepare(nat, xyz, qtab=None):
    """Frames, pair list and the (pairs x NIN) descriptors: 35 pair values,
    then for each atom 56 local values and 4 long-range ones (from the
    charge table qtab: type codes and charges; zero without it); typ per
    pair; the atoms' bonding type codes."""
    nat = np.asarray(nat)
    xyz = np.asarray(xyz, dtype=np.float64)
    nbrs, tree = bonded_neighbours(nat, xyz)
    R = atom_frames(nat, xyz, nbrs, tree)
    D = atom_descriptors(nat, xyz, nbrs, tree, R)
    types = atom_types(nat, nbrs)
    F = np.zeros((len(nat), 4))
    if qtab is not None:
        codes, qs = qtab
        pos = np.searchsorted(codes, types)
        pos = np.minimum(pos, len(codes) - 1)
        q = np.where(codes[pos] == types, qs[pos], 0.0)
        F = field_descriptors(xyz, R, q)
    a, b = pair_list(nat, tree)
    e = np.array([ELEM[int(z)] for z in nat])
    m = len(a)
    same = a == b
    typ = np.where(same, e[a], 4 + 4 * e[a] + e[b])
    X = np.zeros((m, NIN))
    X[np.arange(m), typ] = 1.0
    bond = np.zeros(m)
    nbset = [set(x) for x in nbrs]
    for k in np.flatnonzero(~same):
        bond[k] = 1.0 if int(b[k]) in nbset[int(a[k])] else 0.0
    X[:, 20] = bond
    d = np.einsum("mij,mj->mi", R[a], xyz[b] - xyz[a])
    r = np.linalg.norm(d, axis=1)
    X[:, 21:24] = d / RCUT
    X[:, 24] = r / RCUT
    X[:, 25] = np.where(r > 1e-6, 1.0 / np.maximum(r, 1e-6), 0.0)
    X[:, 26:35] = np.einsum("mij,mkj->mik", R[a], R[b]).reshape(m, 9)
    X[:, 35:35 + NATOM] = D[a]
    X[:, 35 + NATOM:35 + 2 * NATOM] = D[b]
    X[:, 35 + 2 * NATOM:39 + 2 * NATOM] = F[a]
    X[:, 39 + 2 * NATOM:43 + 2 * NATOM] = F[b]
    return {"a": a, "b": b, "R": R, "X": X.astype(np.float32), "typ": typ, "types": types}


def _atom_U(nat, R):
    U = np.tile(np.eye(4), (len(nat), 1, 1))
    heavy = np.asarray(nat) != 1
    U[heavy, 1:, 1:] = np.transpose(R[heavy], (0, 2, 1))
    return U


def blocks_from_density(system, p, prep):
    """Training targets: each pair's density block in the atom frames."""
    a, b = prep["a"], prep["b"]
    U = _atom_U(system.nat, prep["R"])
    idx, valid, _ = G._slots(system, a, b)
    blk = np.where(valid, p[idx], 0.0)
    same = a == b
    blk[same] = np.tril(blk[same]) + np.transpose(np.tril(blk[same], -1), (0, 2, 1))
    return np.einsum("mxi,mxy,myj->mij", U[a], blk, U[b]).reshape(len(a), NOUT)


def assemble(system, y, prep, renormalize=True):
    """Packed starting density from the pair blocks (frames of A and B)."""
    a, b = prep["a"], prep["b"]
    U = _atom_U(system.nat, prep["R"])
    L_ = np.asarray(y, dtype=np.float64).reshape(len(a), 4, 4).copy()
    same = a == b
    L_[same] = 0.5 * (L_[same] + np.transpose(L_[same], (0, 2, 1)))
    blk = np.einsum("mix,mxy,mjy->mij", U[a], L_, U[b])
    idx, valid, _ = G._slots(system, a, b)
    p = np.zeros(system.mpack)
    p[idx[valid]] = blk[valid]
    if renormalize:
        frag = geometry.fragments(system.nat, system.xyz)
        no = (system.nlast - system.nfirst + 1).astype(np.int64)
        orb_frag = np.repeat(frag, no)
        orbs = np.arange(system.norbs)
        d_idx = orbs * (orbs + 1) // 2 + orbs
        want = np.bincount(frag, weights=[VALENCE[int(z)] for z in system.nat])
        tot = np.bincount(orb_frag, weights=p[d_idx], minlength=len(want))
        sc = np.where(tot > 1e-6, want / np.maximum(tot, 1e-6), 1.0)
        p[d_idx] *= sc[orb_frag]
    return p


class HcnoGuessModel:
    """Trained network, input normalization, per-pair-type output
    normalization (y_mean, y_std: 20 x 16) and the charge table of the
    long-range descriptors (sorted type codes, charges)."""

    def __init__(self, params, x_mean, x_std, y_mean, y_std, q_codes, q_values):
        self.params = np.asarray(params, dtype=np.float32)
        self.x_mean, self.x_std = x_mean, x_std
        self.y_mean, self.y_std = y_mean, y_std
        self.qtab = (np.asarray(q_codes, dtype=np.int64), np.asarray(q_values, dtype=np.float64))
        self._mlp = None

    @classmethod
    def load(cls, path=DEFAULT_MODEL):
        z = np.load(path, allow_pickle=False)
        return cls(z["params"], z["x_mean"], z["x_std"], z["y_mean"], z["y_std"], z["q_codes"], z["q_values"])

    def save(self, path=DEFAULT_MODEL):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params, x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean,
                 y_std=self.y_std, q_codes=self.qtab[0], q_values=self.qtab[1])

    def mlp(self):
        if self._mlp is None:
            self._mlp = G.MLP(self.params, dims=layer_dims())
        return self._mlp

    def predict_blocks(self, prep):
        xn = ((prep["X"] - self.x_mean) / self.x_std).astype(np.float32)
        t = prep["typ"]
        return self.mlp().forward(xn) * self.y_std[t] + self.y_mean[t]

    def predict_density(self, system):
        """Packed starting density, or None when an element is not H, C, N, O."""
        if not supported(system.nat):
            return None
        prep = prepare(system.nat, system.xyz, self.qtab)
        return assemble(system, self.predict_blocks(prep), prep)


def load_default():
    if not os.path.exists(DEFAULT_MODEL):
        return None
    return HcnoGuessModel.load(DEFAULT_MODEL)
