"""Learned starting density for water clusters.

A small MLP (3 x 256 hidden units, BF16 weights, FP32 accumulation; kernels in
k_mlp.hip) maps an atom pair -- its element-pair type plus geometry and radial
environment descriptors within 4 A -- to that pair's density-matrix block:

    self O (4x4), self H (1x1), O-H (4x1), H-O (1x4), H-H (1x1), O-O (4x4),

for the atom itself, its own molecule and every atom closer than 4 A.  Blocks
are predicted in the frames of the molecules the two atoms belong to, so the
prediction rotates with the system; they are rotated to the molecular frame,
assembled into a packed starting density (blocks beyond 4 A are zero) and
renormalized to 8 valence electrons per molecule.  The SCF starts from it as
the reference starts from an OLDENS density (iter.F90:151-159).  When the
system is not made of water molecules, or no trained model is present, the
standard diagonal guess (moldat.F90:731-790) is used instead.
"""
import ctypes
import os

import numpy as np
from scipy.spatial import cKDTree

from . import lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "water_guess.npz")

NIN = 96
NHID = 256
NLAYERS = 3
NOUT = 16
RCUT = 4.0
MUS = np.linspace(1.0, 4.0, 6)
ETA = 3.0
# environment field: point charges of the isolated PM7 water molecule
# (the reference's API test values, tests/mopac_api_test.F90:92-94)
Q_O, Q_H = -0.644520, 0.322260
RFIELD = 9.0
SHELLS = ((1.4, 2.6), (2.6, 4.0))

# pair types
SELF_O, SELF_H, INTRA_OH, INTRA_HH, INTER_OO, INTER_OH, INTER_HO, INTER_HH = range(8)


def find_waters(nat, xyz):
    """(mol_of_atom, waters) when every atom belongs to a water molecule, else None.
    waters[k] = (iO, iH1, iH2), the two H in atom order."""
    nat = np.asarray(nat)
    xyz = np.asarray(xyz, dtype=np.float64)
    o_idx = np.where(nat == 8)[0]
    h_idx = np.where(nat == 1)[0]
    if len(o_idx) == 0 or len(o_idx) * 3 != len(nat) or len(h_idx) != 2 * len(o_idx):
        return None
    d, owner = cKDTree(xyz[o_idx]).query(xyz[h_idx])
    if np.any(d > 1.25) or np.any(np.bincount(owner, minlength=len(o_idx)) != 2):
        return None
    hs = h_idx[np.argsort(owner, kind="stable")].reshape(-1, 2)
    w = np.column_stack([o_idx, hs])
    mol = np.empty(len(nat), dtype=np.int64)
    mol[w.ravel()] = np.repeat(np.arange(len(o_idx)), 3)
    return mol, [tuple(int(v) for v in row) for row in w]


def frames(xyz, waters):
    """(nw, 3, 3): rows e1 (bisector), e2 (toward H1, in plane), e3 (normal)."""
    w = np.array(waters)
    o, h1, h2 = xyz[w[:, 0]], xyz[w[:, 1]], xyz[w[:, 2]]
    b = (h1 - o) + (h2 - o)
    e1 = b / np.linalg.norm(b, axis=1, keepdims=True)
    d = h1 - h2
    d = d - np.sum(d * e1, axis=1, keepdims=True) * e1
    e2 = d / np.linalg.norm(d, axis=1, keepdims=True)
    e3 = np.cross(e1, e2)
    return np.stack([e1, e2, e3], axis=1)


def _radial(nat, xyz, mol, tree):
    """Per atom: Gaussian radial densities of O and H neighbours of other molecules."""
    n = len(nat)
    feat = np.zeros((n, 2 * len(MUS)))
    pairs = tree.query_pairs(RCUT, output_type="ndarray")
    if len(pairs):
        i, j = pairs[:, 0], pairs[:, 1]
        keep = mol[i] != mol[j]
        i, j = i[keep], j[keep]
        r = np.linalg.norm(xyz[i] - xyz[j], axis=1)
        fc = 0.5 * (np.cos(np.pi * r / RCUT) + 1.0)
        g = np.exp(-ETA * (r[:, None] - MUS[None, :]) ** 2) * fc[:, None]
        for a, b in ((i, j), (j, i)):
            col = np.where(nat[b] == 8, 0, len(MUS))
            for m in range(len(MUS)):
                np.add.at(feat, (a, col + m), g[:, m])
    return feat


def _environment(nat, xyz, mol, R):
    """Per atom, in the frame of its own molecule: the electrostatic potential,
    field and field gradient (5) of the other molecules' point charges within
    RFIELD (smoothly switched off over the last 1 A), and the unit-vector sums
    toward O and H neighbours of other molecules in two distance shells."""
    nat = np.asarray(nat)
    n = len(nat)
    q = np.where(nat == 8, Q_O, Q_H)
    ncol = 9 + 3 * 2 * len(SHELLS)
    feat = np.zeros((n, ncol))
    pairs = cKDTree(xyz).query_pairs(RFIELD, output_type="ndarray")
    if not len(pairs):
        return feat
    i, j = pairs[:, 0], pairs[:, 1]
    keep = mol[i] != mol[j]
    i, j = i[keep], j[keep]
    a = np.concatenate([i, j])  # both directions at once
    b = np.concatenate([j, i])
    d = xyz[b] - xyz[a]
    r = np.sqrt(np.einsum("mi,mi->m", d, d))
    u = np.einsum("mij,mj->mi", R[mol[a]], d / r[:, None])
    sw = np.where(r < RFIELD - 1.0, 1.0, 0.5 * (1.0 + np.cos(np.pi * (r - (RFIELD - 1.0)))))
    qs = q[b] * sw
    cols = np.empty((len(a), ncol))
    cols[:, 0] = qs / r
    cols[:, 1:4] = (qs / r ** 2)[:, None] * u
    g = qs / r ** 3
    cols[:, 4] = g * (3 * u[:, 0] ** 2 - 1)
    cols[:, 5] = g * (3 * u[:, 1] ** 2 - 1)
    cols[:, 6] = g * 3 * u[:, 0] * u[:, 1]
    cols[:, 7] = g * 3 * u[:, 0] * u[:, 2]
    cols[:, 8] = g * 3 * u[:, 1] * u[:, 2]
    cols[:, 9:] = 0.0
    for si, (lo, hi) in enumerate(SHELLS):
        in_shell = (r >= lo) & (r < hi)
        for ei, el in enumerate((8, 1)):
            m = in_shell & (nat[b] == el)
            c0 = 9 + 3 * (2 * si + ei)
            cols[m, c0:c0 + 3] = u[m]
    # scatter-add by atom: one bincount per column (np.add.at is far slower)
    for k in range(ncol):
        feat[:, k] = np.bincount(a, weights=cols[:, k], minlength=n)
    return feat


def pair_list(nat, xyz, mol):
    """All (A, B) with A >= B: self pairs, same-molecule pairs, pairs within RCUT."""
    tree = cKDTree(xyz)
    pairs = tree.query_pairs(RCUT, output_type="ndarray")
    n = len(nat)
    a = np.concatenate([np.arange(n), np.maximum(pairs[:, 0], pairs[:, 1]) if len(pairs) else []]).astype(np.int64)
    b = np.concatenate([np.arange(n), np.minimum(pairs[:, 0], pairs[:, 1]) if len(pairs) else []]).astype(np.int64)
    # same-molecule pairs beyond RCUT cannot happen for water (H-H ~ 1.5 A)
    return a, b, tree


def features(nat, xyz, mol, waters, a, b, tree):
    nat = np.asarray(nat)
    R = frames(xyz, waters)
    w = np.array(waters)
    rad = _radial(nat, xyz, mol, tree)
    m = len(a)
    X = np.zeros((m, NIN), dtype=np.float64)
    same = a == b
    intra = (mol[a] == mol[b]) & ~same
    ta, tb = nat[a] == 8, nat[b] == 8
    typ = np.where(same & ta, SELF_O, 0)
    typ = np.where(same & ~ta, SELF_H, typ)
    typ = np.where(intra & (ta | tb), INTRA_OH, typ)
    typ = np.where(intra & ~ta & ~tb, INTRA_HH, typ)
    inter = ~same & ~intra
    typ = np.where(inter & ta & tb, INTER_OO, typ)
    typ = np.where(inter & ta & ~tb, INTER_OH, typ)
    typ = np.where(inter & ~ta & tb, INTER_HO, typ)
    typ = np.where(inter & ~ta & ~tb, INTER_HH, typ)
    X[np.arange(m), typ] = 1.0
    Ra, Rb = R[mol[a]], R[mol[b]]
    d = np.einsum("mij,mj->mi", Ra, xyz[b] - xyz[a])
    r = np.linalg.norm(d, axis=1)
    X[:, 8:11] = d / 4.0
    X[:, 11] = r / 4.0
    X[:, 12] = np.where(r > 1e-6, 1.0 / np.maximum(r, 1e-6), 0.0)
    X[:, 13:22] = np.einsum("mij,mkj->mik", Ra, Rb).reshape(m, 9)
    X[:, 22:34] = rad[a]
    X[:, 34:46] = rad[b]
    o, h1, h2 = xyz[w[:, 0]], xyz[w[:, 1]], xyz[w[:, 2]]
    r1 = np.linalg.norm(h1 - o, axis=1)
    r2 = np.linalg.norm(h2 - o, axis=1)
    ca = np.sum((h1 - o) * (h2 - o), axis=1) / (r1 * r2)
    geo = np.stack([r1 - 0.96, r2 - 0.96, ca - np.cos(np.radians(104.5))], axis=1)
    X[:, 46:49] = geo[mol[a]]
    X[:, 49:52] = geo[mol[b]]
    # which hydrogen of its molecule each atom is (0 for O, +-1 for H1/H2)
    side = np.zeros(len(nat))
    side[w[:, 1]] = 1.0
    side[w[:, 2]] = -1.0
    X[:, 52] = side[a]
    X[:, 53] = side[b]
    env = _environment(nat, xyz, mol, R)
    X[:, 54:75] = env[a]
    X[:, 75:96] = env[b]
    return X.astype(np.float32), typ


def _rot(R):
    """Orbital rotation of one sp atom: molecular coefficients = U @ local."""
    U = np.eye(4)
    U[1:, 1:] = R.T
    return U


def _atom_U(nat, R, mol):
    U = np.tile(np.eye(4), (len(nat), 1, 1))
    o = np.asarray(nat) == 8
    U[o, 1:, 1:] = np.transpose(R[mol[o]], (0, 2, 1))
    return U


def _slots(system, a, b):
    """Packed index of every (x, y) slot of each pair block, and its validity."""
    nf = (system.nfirst - 1).astype(np.int64)
    no = (system.nlast - system.nfirst + 1).astype(np.int64)
    x = np.arange(4)
    r = nf[a][:, None, None] + x[None, :, None]
    c = nf[b][:, None, None] + x[None, None, :]
    valid = (x[None, :, None] < no[a][:, None, None]) & (x[None, None, :] < no[b][:, None, None])
    valid &= ~((a == b)[:, None, None] & (c > r))
    hi = np.maximum(r, c)
    lo = np.minimum(r, c)
    idx = hi * (hi + 1) // 2 + lo
    return np.where(valid, idx, 0), valid, (hi == lo) & valid


def blocks_from_density(system, p, a, b, mol, waters):
    """Targets: each pair's density block in the molecule frames (16 values)."""
    U = _atom_U(system.nat, frames(system.xyz, waters), mol)
    idx, valid, _ = _slots(system, a, b)
    blk = np.where(valid, p[idx], 0.0)
    same = a == b
    # self blocks: fill the upper triangle from the lower one
    blk[same] = np.tril(blk[same]) + np.transpose(np.tril(blk[same], -1), (0, 2, 1))
    Lb = np.einsum("mxi,mxy,myj->mij", U[a], blk, U[b])
    return Lb.reshape(len(a), NOUT)


def assemble(system, y, a, b, mol, waters):
    """Packed starting density from predicted pair blocks."""
    U = _atom_U(system.nat, frames(system.xyz, waters), mol)
    L_ = y.reshape(len(a), 4, 4).astype(np.float64)
    same = a == b
    L_[same] = 0.5 * (L_[same] + np.transpose(L_[same], (0, 2, 1)))
    blk = np.einsum("mix,mxy,mjy->mij", U[a], L_, U[b])
    idx, valid, dg = _slots(system, a, b)
    p = np.zeros(system.mpack)
    p[idx[valid]] = blk[valid]
    # 8 valence electrons per molecule: rescale each molecule's diagonal
    nf = (system.nfirst - 1).astype(np.int64)
    no = (system.nlast - system.nfirst + 1).astype(np.int64)
    orb_mol = np.repeat(mol, no)
    orbs = np.arange(system.norbs)
    d_idx = orbs * (orbs + 1) // 2 + orbs
    tot = np.bincount(orb_mol, weights=p[d_idx], minlength=len(waters))
    sc = np.where(tot > 1e-6, 8.0 / np.maximum(tot, 1e-6), 1.0)
    p[d_idx] *= sc[orb_mol]
    return p


# ---- the network (GPU kernels in libmopachip) --------------------------------

def layer_dims():
    return [NIN] + [NHID] * NLAYERS + [NOUT]


def n_params(dims=None):
    d = dims or layer_dims()
    return sum(d[i] * d[i + 1] + d[i + 1] for i in range(len(d) - 1))


def init_params(seed=0, dims=None):
    rng = np.random.default_rng(seed)
    d = dims or layer_dims()
    parts = []
    for i in range(len(d) - 1):
        w = rng.standard_normal((d[i], d[i + 1])) * np.sqrt(1.0 / d[i])
        parts += [w.ravel(), np.zeros(d[i + 1])]
    return np.concatenate(parts).astype(np.float32)


class MLP:
    """Handle on the device MLP (mhip_mlp_*)."""

    def __init__(self, params=None, seed=0, dims=None):
        self.lib = L.load()
        self.h = ctypes.c_void_p()
        self.dims = list(dims or layer_dims())
        d = np.array(self.dims, dtype=np.int32)
        L.check(self.lib.mhip_mlp_create(len(d) - 1, L.ptr(d), ctypes.byref(self.h)))
        self.set_params(init_params(seed, self.dims) if params is None else params)

    def close(self):
        if self.h:
            self.lib.mhip_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != n_params(self.dims):
            raise ValueError(f"MLP parameter vector has {params.size} values, the network {n_params(self.dims)}")
        L.check(self.lib.mhip_mlp_set_params(self.h, L.ptr(params)))

    def get_params(self):
        out = np.zeros(n_params(self.dims), dtype=np.float32)
        L.check(self.lib.mhip_mlp_get_params(self.h, L.ptr(out)))
        return out

    def _check_x(self, x):
        if x.ndim != 2 or x.shape[1] != self.dims[0] or x.shape[0] == 0:
            raise ValueError(f"MLP input must be (batch > 0, {self.dims[0]}), got {x.shape}")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        self._check_x(x)
        y = np.zeros((x.shape[0], self.dims[-1]), dtype=np.float32)
        L.check(self.lib.mhip_mlp_forward(self.h, x.shape[0], L.ptr(x), L.ptr(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._check_x(x)
        if t.shape != (x.shape[0], self.dims[-1]):
            raise ValueError(f"MLP targets must be ({x.shape[0]}, {self.dims[-1]}), got {t.shape}")
        g = np.zeros(n_params(self.dims), dtype=np.float32)
        loss = ctypes.c_float()
        L.check(self.lib.mhip_mlp_gradients(self.h, x.shape[0], L.ptr(x), L.ptr(t), L.ptr(g),
                                            ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._check_x(x)
        if t.shape != (x.shape[0], self.dims[-1]):
            raise ValueError(f"MLP targets must be ({x.shape[0]}, {self.dims[-1]}), got {t.shape}")
        loss = ctypes.c_float()
        L.check(self.lib.mhip_mlp_train_step(self.h, x.shape[0], L.ptr(x), L.ptr(t), ctypes.c_float(lr),
                                             ctypes.byref(loss)))
        return loss.value


class GuessModel:
    """Trained network plus input normalization and per-pair-type output
    normalization (y_mean, y_std: 8 x 16)."""

    def __init__(self, params, x_mean, x_std, y_mean, y_std):
        self.params = np.asarray(params, dtype=np.float32)
        self.x_mean, self.x_std = x_mean, x_std
        self.y_mean, self.y_std = y_mean, y_std
        self._mlp = None

    @classmethod
    def load(cls, path=DEFAULT_MODEL):
        z = np.load(path, allow_pickle=False)
        return cls(z["params"], z["x_mean"], z["x_std"], z["y_mean"], z["y_std"])

    def save(self, path=DEFAULT_MODEL):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params, x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean,
                 y_std=self.y_std)

    def mlp(self):
        if self._mlp is None:
            self._mlp = MLP(self.params)
        return self._mlp

    def to_device(self, ctx, system, want_xy=False):
        """Build the starting density on the device (mhip_guess_water); the
        next ctx.iter(start_from_guess=True) starts from it.  Returns None
        (standard guess) when the system is not all water, else a dict with
        the pair list and, with want_xy, the normalized descriptors and raw
        network outputs."""
        fw = find_waters(system.nat, system.xyz)
        if fw is None:
            return None
        mol, waters = fw
        a, b, _ = pair_list(system.nat, system.xyz, mol)
        nat = np.ascontiguousarray(system.nat, dtype=np.int32)
        xyz = np.ascontiguousarray(system.xyz, dtype=np.float64)
        w = np.ascontiguousarray(np.array(waters), dtype=np.int32)
        pa = np.ascontiguousarray(a, dtype=np.int32)
        pb = np.ascontiguousarray(b, dtype=np.int32)
        norm = [np.ascontiguousarray(v, dtype=np.float32) for v in (self.x_mean, self.x_std)]
        norm += [np.ascontiguousarray(np.broadcast_to(v, (8, NOUT)), dtype=np.float32)
                 for v in (self.y_mean, self.y_std)]
        model = L.GuessModelC(*[v.ctypes.data for v in norm])
        x = np.zeros((len(a), NIN), dtype=np.float32) if want_xy else None
        y = np.zeros((len(a), NOUT), dtype=np.float32) if want_xy else None
        L.check(ctx.lib.mhip_guess_water(ctx.h, self.mlp().h, ctypes.byref(model), L.ptr(nat), L.ptr(xyz),
                                         len(waters), L.ptr(w), len(a), L.ptr(pa), L.ptr(pb), L.ptr(x),
                                         L.ptr(y)))
        return {"a": a, "b": b, "mol": mol, "waters": waters, "x": x, "y": y}

    def predict_density(self, system, ctx=None):
        """Packed starting density computed on the host (descriptors and
        assembly in numpy, network on the device), or None (standard guess)
        when not all water.  The device path is to_device."""
        fw = find_waters(system.nat, system.xyz)
        if fw is None:
            return None
        mol, waters = fw
        a, b, tree = pair_list(system.nat, system.xyz, mol)
        X, _ = features(system.nat, system.xyz, mol, waters, a, b, tree)
        xn = ((X - self.x_mean) / self.x_std).astype(np.float32)
        typ = np.argmax(X[:, :8], axis=1)
        ym = self.y_mean[typ] if self.y_mean.ndim == 2 else self.y_mean
        ys = self.y_std[typ] if self.y_std.ndim == 2 else self.y_std
        y = self.mlp().forward(xn) * ys + ym
        return assemble(system, y, a, b, mol, waters)


def load_default():
    if not os.path.exists(DEFAULT_MODEL):
        return None
    return GuessModel.load(DEFAULT_MODEL)
