"""Seeded synthetic geometries for the SCF workloads (Angstrom).

- water_cluster: O on a cubic lattice (default 3.1 A) with uniform +-0.2 A
  jitter per coordinate; each H2O has O-H 0.96 A, HOH 104.5 deg and a
  Haar-uniform random orientation (unit quaternion).
- alkane: zig-zag CnH2n+2 chain, C-C 1.54 A, C-H 1.09 A, tetrahedral angles,
  random torsions about the chain.
- glycine_helix: poly-glycine alpha helix (phi -57, psi -47, omega 180) with
  ideal bond lengths and seeded +-5 deg torsion noise, capped H / OH.
Generators use numpy's PCG64 with the stated seed, so geometries are
reproducible across machines.
"""
import numpy as np

Z = {"H": 1, "C": 6, "N": 7, "O": 8}


def _quat_rot(rng):
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def lattice_shape(k):
    """Near-cubic lattice dims holding k molecules (4x4x2 for 32, 8x8x8 for 512)."""
    best = None
    for a in range(1, k + 1):
        for b in range(a, k + 1):
            if a * b > k:
                break
            c = -(-k // (a * b))
            cand = (a * b * c - k, c - a, (c, b, a))
            if best is None or cand < best:
                best = cand
    return best[2]


def water_cluster(k, seed, spacing=3.1, jitter=0.2, shape=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = shape if shape is not None else lattice_shape(k)
    half = np.radians(104.5) / 2
    local = np.array([[0.0, 0.0, 0.0],
                      [0.96 * np.sin(half), 0.96 * np.cos(half), 0.0],
                      [-0.96 * np.sin(half), 0.96 * np.cos(half), 0.0]])
    nat, xyz = [], []
    m = 0
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                if m == k:
                    break
                o = np.array([ix, iy, iz], dtype=float) * spacing + rng.uniform(-jitter, jitter, 3)
                R = _quat_rot(rng)
                for zat, p in zip((8, 1, 1), local):
                    nat.append(zat)
                    xyz.append(o + R @ p)
                m += 1
    return np.array(nat, dtype=np.int32), np.array(xyz, dtype=np.float64)


def _place(a, b, c, bond, angle, torsion):
    """Atom d bonded to c, with |cd| = bond, angle bcd, dihedral abcd (radians)."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * np.cos(angle), bond * np.sin(angle) * np.cos(torsion),
                   bond * np.sin(angle) * np.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def alkane(n, seed):
    """CnH2n+2, zig-zag backbone with random torsions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tet = np.radians(109.47)
    cs = [np.zeros(3)]
    if n >= 2:
        cs.append(np.array([1.54, 0.0, 0.0]))
    if n >= 3:
        cs.append(cs[1] + 1.54 * np.array([np.cos(np.pi - tet), np.sin(np.pi - tet), 0.0]))
    for i in range(3, n):
        tor = np.radians(180.0 + rng.uniform(-60.0, 60.0))
        cs.append(_place(cs[i - 3], cs[i - 2], cs[i - 1], 1.54, tet, tor))
    nat = [6] * n
    xyz = list(cs)
    for i in range(n):
        nb = [cs[j] for j in (i - 1, i + 1) if 0 <= j < n]
        nh = 4 - len(nb)
        c = cs[i]
        if len(nb) == 0:
            dirs = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) / np.sqrt(3)
        elif len(nb) == 1:
            u = (c - nb[0]) / np.linalg.norm(c - nb[0])
            t = np.cross(u, [0.0, 0.0, 1.0] if abs(u[2]) < 0.9 else [1.0, 0.0, 0.0])
            t /= np.linalg.norm(t)
            w = np.cross(u, t)
            ang = tet
            phi0 = rng.uniform(0, 2 * np.pi)
            dirs = [(-np.cos(ang)) * u + np.sin(ang) * (np.cos(phi0 + k * 2 * np.pi / 3) * t +
                                                        np.sin(phi0 + k * 2 * np.pi / 3) * w)
                    for k in range(3)]
        else:
            u1 = (nb[0] - c) / np.linalg.norm(nb[0] - c)
            u2 = (nb[1] - c) / np.linalg.norm(nb[1] - c)
            bis = -(u1 + u2)
            bis /= np.linalg.norm(bis)
            perp = np.cross(u1, u2)
            perp /= np.linalg.norm(perp)
            dirs = [np.cos(tet / 2) * bis + s * np.sin(tet / 2) * perp for s in (1.0, -1.0)]
        for k in range(nh):
            nat.append(1)
            xyz.append(c + 1.09 * np.asarray(dirs[k]))
    return np.array(nat, dtype=np.int32), np.array(xyz, dtype=np.float64)


def glycine_helix(nres, seed):
    """Poly-glycine alpha helix, N-capped with H and C-capped with OH."""
    rng = np.random.Generator(np.random.PCG64(seed))
    phi0, psi0, omg = -57.0, -47.0, 180.0
    nat, xyz = [], []
    # seed the first three backbone atoms (N, CA, C) with ideal geometry
    N = np.array([0.0, 0.0, 0.0])
    CA = np.array([1.458, 0.0, 0.0])
    ang = np.radians(111.0)
    C = CA + 1.525 * np.array([-np.cos(ang), np.sin(ang), 0.0])
    backbone = []
    for r in range(nres):
        if r > 0:
            prevN, prevCA, prevC = backbone[-1]
            psi = np.radians(psi0 + rng.uniform(-5, 5))
            om = np.radians(omg + rng.uniform(-5, 5))
            phi = np.radians(phi0 + rng.uniform(-5, 5))
            N = _place(prevN, prevCA, prevC, 1.329, np.radians(116.2), psi)
            CA = _place(prevCA, prevC, N, 1.458, np.radians(121.7), om)
            C = _place(prevC, N, CA, 1.525, np.radians(111.0), phi)
        backbone.append((N, CA, C))
    for r, (N, CA, C) in enumerate(backbone):
        nxt = backbone[r + 1][0] if r + 1 < nres else None
        prevC = backbone[r - 1][2] if r > 0 else None
        nat += [7, 6, 6]
        xyz += [N, CA, C]
        # carbonyl O in the peptide plane, opposite the next N
        ref = nxt if nxt is not None else CA + (C - CA) * 2.0 + np.array([0.3, 0.2, 0.1])
        u1 = (CA - C) / np.linalg.norm(CA - C)
        u2 = (ref - C) / np.linalg.norm(ref - C)
        o_dir = -(u1 + u2)
        o_dir /= np.linalg.norm(o_dir)
        nat.append(8)
        xyz.append(C + 1.231 * o_dir)
        # amide H (or N-terminal cap H)
        if prevC is not None:
            v1 = (prevC - N) / np.linalg.norm(prevC - N)
            v2 = (CA - N) / np.linalg.norm(CA - N)
            h_dir = -(v1 + v2)
            h_dir /= np.linalg.norm(h_dir)
            nat.append(1)
            xyz.append(N + 1.01 * h_dir)
        else:
            for tor in (60.0, 180.0):
                nat.append(1)
                xyz.append(_place(C, CA, N, 1.01, np.radians(109.5), np.radians(tor)))
        # two alpha hydrogens
        for tor in (120.0, -120.0):
            nat.append(1)
            xyz.append(_place(N, C, CA, 1.09, np.radians(109.5), np.radians(tor)))
        if nxt is None:
            # C-terminal OH
            oh = _place(N, CA, C, 1.36, np.radians(111.0), np.radians(180.0))
            nat.append(8)
            xyz.append(oh)
            nat.append(1)
            xyz.append(_place(CA, C, oh, 0.96, np.radians(107.0), np.radians(0.0)))
    return np.array(nat, dtype=np.int32), np.array(xyz, dtype=np.float64)


COVALENT_RADIUS = {1: 0.31, 6: 0.76, 7: 0.71, 8: 0.66}


def fragments(nat, xyz, scale=1.2):
    """Atom -> fragment id: connected components of the covalent bond graph
    (bonded when closer than scale x the sum of covalent radii)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial import cKDTree
    nat = np.asarray(nat)
    xyz = np.asarray(xyz, dtype=np.float64)
    rad = np.array([COVALENT_RADIUS[int(z)] for z in nat])
    pairs = cKDTree(xyz).query_pairs(2.0 * scale * rad.max(), output_type="ndarray")
    if len(pairs):
        d = np.linalg.norm(xyz[pairs[:, 0]] - xyz[pairs[:, 1]], axis=1)
        pairs = pairs[d < scale * (rad[pairs[:, 0]] + rad[pairs[:, 1]])]
    n = len(nat)
    g = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n)) if len(pairs) else \
        coo_matrix((n, n))
    _, labels = connected_components(g, directed=False)
    return labels.astype(np.int32)
