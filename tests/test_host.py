"""CPU tests of the host side: the C-ABI library loads and exports what the
header declares, the product package never reaches for the oracle, the
Hamiltonian producer's structure, the learned-guess host logic, the MLP
reference's gradients, and the bench's one-GPU contract."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from mopac_gpu_amd import geometry, guess, hamiltonian
from mopac_gpu_amd import lib as L
from oracle import mlp_ref, oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- the C ABI ----------------------------------------------------------------

def test_library_exports_every_header_symbol():
    names = L.exported_symbols()
    assert len(names) >= 29
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = L.load()  # loads without a GPU; every symbol resolves through ctypes
    for n in names:
        assert getattr(lib, n) is not None
    assert set(names) <= set(L.SIGNATURES)


def test_library_carries_gfx950_code_only():
    blob = open(L.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-f]+)", blob))
    assert targets == {b"gfx950"}, targets


def test_product_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "mopac_gpu_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, fn), encoding="utf-8").read()
                assert not re.search(r"^\s*(from|import)\s+oracle", text, re.M), fn
                assert "liboracle" not in text and "oracle_scf" not in text, fn


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


# ---- the Hamiltonian producer -----------------------------------------------

def test_hamiltonian_sizes_and_translation_invariance():
    nat, xyz = geometry.glycine_helix(2, 1)
    s0 = hamiltonian.build(nat, xyz)
    s1 = hamiltonian.build(nat, xyz + np.array([10.0, -3.0, 7.5]))
    heavy = int(np.sum(np.asarray(nat) > 1))
    assert s0.norbs == 4 * heavy + (len(nat) - heavy)
    assert s0.mpack == s0.norbs * (s0.norbs + 1) // 2
    assert np.max(np.abs(s0.h - s1.h)) < 1e-10
    assert np.max(np.abs(s0.w - s1.w)) < 1e-10
    assert abs(s0.enuclr - s1.enuclr) < 1e-9


def test_water_cluster_geometry():
    nat, xyz = geometry.water_cluster(64, 3)
    assert len(nat) == 192 and np.sum(nat == 8) == 64
    fw = guess.find_waters(nat, xyz)
    assert fw is not None
    _, waters = fw
    w = np.array(waters)
    roh = np.linalg.norm(xyz[w[:, 1:]] - xyz[w[:, :1]], axis=2)
    assert np.all(np.abs(roh - 0.96) < 0.1)
    o = xyz[w[:, 0]]
    d = np.linalg.norm(o[:, None] - o[None], axis=2) + 1e3 * np.eye(64)
    assert d.min() > 2.3


# ---- learned-guess host logic -------------------------------------------------

def _rotated(xyz, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return xyz @ q.T + rng.standard_normal(3)


def test_pair_list_matches_brute_force():
    nat, xyz = geometry.water_cluster(27, 5)
    mol, _ = guess.find_waters(nat, xyz)
    a, b, _ = guess.pair_list(nat, xyz, mol)
    got = set(zip(a.tolist(), b.tolist()))
    n = len(nat)
    want = {(i, i) for i in range(n)}
    for i in range(n):
        for j in range(i):
            if np.linalg.norm(xyz[i] - xyz[j]) < guess.RCUT or mol[i] == mol[j]:
                want.add((i, j))
    assert got == want


def test_features_are_rotation_invariant():
    nat, xyz = geometry.water_cluster(16, 2)
    x1 = _rotated(xyz, 9)
    out = []
    for c in (xyz, x1):
        mol, waters = guess.find_waters(nat, c)
        a, b, tree = guess.pair_list(nat, c, mol)
        order = np.lexsort((b, a))
        X, typ = guess.features(nat, c, mol, waters, a, b, tree)
        out.append(X[order])
    assert np.max(np.abs(out[0] - out[1])) < 2e-5


def test_blocks_roundtrip_and_equivariance():
    """Targets computed from a converged density rotate with the molecule, and
    assembling them back reproduces the density on every covered element."""
    nat, xyz = geometry.water_cluster(6, 4)
    res = []
    for c in (xyz, _rotated(xyz, 3)):
        s = hamiltonian.build(nat, c)
        o = oracle.iterate(s, scfcrt=1e-9)
        mol, waters = guess.find_waters(nat, c)
        a, b, _ = guess.pair_list(nat, c, mol)
        order = np.lexsort((b, a))
        y = guess.blocks_from_density(s, o["p"], a, b, mol, waters)
        p = guess.assemble(s, y, a, b, mol, waters)
        d = np.array([i * (i + 1) // 2 + i for i in range(s.norbs)])
        covered = np.zeros(s.mpack, bool)
        idx, valid, _ = guess._slots(s, a, b)
        covered[idx[valid]] = True
        off = covered.copy()
        off[d] = False
        assert np.max(np.abs(p[off] - o["p"][off])) < 1e-12
        assert abs(p[d].sum() - 8 * len(waters)) < 1e-9
        res.append(y[order])
    assert np.max(np.abs(res[0] - res[1])) < 1e-6


def test_mlp_reference_gradient_matches_finite_differences():
    """The backward pass of the CPU MLP reference, with BF16 rounding switched
    off and FP64 parameters, against central differences of its own loss."""
    dims = [5, 7, 6, 3]
    rng = np.random.default_rng(0)
    nparam = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))
    params = rng.standard_normal(nparam) * 0.5
    x = rng.standard_normal((9, dims[0]))
    t = rng.standard_normal((9, dims[-1]))
    ident = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731
    loss0, g = mlp_ref.loss_and_grad(params, dims, x, t, rnd=ident, pdtype=np.float64)
    eps = 1e-6
    fd = np.array([(mlp_ref.loss_and_grad(params + eps * e, dims, x, t, ident, np.float64)[0]
                    - mlp_ref.loss_and_grad(params - eps * e, dims, x, t, ident, np.float64)[0]) / (2 * eps)
                   for e in np.eye(nparam)])
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))
    # and the default (BF16) forward differs from the FP64 one only by rounding
    y16 = mlp_ref.forward(params.astype(np.float32), dims, x)[-1]
    y64 = mlp_ref.forward(params, dims, x, ident, np.float64)[-1]
    assert 0 < np.max(np.abs(y16 - y64)) < 5e-2 * np.max(np.abs(y64))


def test_fortran_interface_structs_match_c_layout(tmp_path):
    """The bind(C) types of the Fortran host module have the C structs' sizes."""
    import ctypes
    from mopac_gpu_amd import build as B
    if not os.path.exists(B.FC):
        pytest.skip("no Fortran compiler")
    src = tmp_path / "sz.F90"
    src.write_text("program sz\n  use iso_c_binding\n  use mopac_hip\n  type(mhip_scf_opts) :: o\n"
                   "  type(mhip_scf_result) :: r\n  print *, c_sizeof(o), c_sizeof(r)\nend program sz\n")
    mod = os.path.join(ROOT, "mopac_gpu_amd", "_build", "fortran")
    exe = tmp_path / "sz"
    subprocess.run([B.FC, "-I", mod, str(src), os.path.join(mod, "mopac_hip.o"), "-L", B.HERE, "-lmopachip",
                    "-Wl,-rpath," + B.HERE, "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(L.ScfOpts), ctypes.sizeof(L.ScfResult)]


def test_batch_workload_spans_the_configured_sizes():
    from mopac_gpu_amd import batch as B
    mols = B.workload(200, seed=4)
    sizes = []
    for nat, xyz in mols:
        nat = np.asarray(nat)
        heavy = int(np.sum(nat > 1))
        sizes.append(4 * heavy + (len(nat) - heavy))
    sizes = np.array(sizes)
    assert sizes.min() >= 6 and sizes.max() <= 180
    alk = sizes[0::2]
    assert np.all((alk - 2) % 6 == 0)            # CnH2n+2: 6n + 2 basis functions
    wat = sizes[1::2]
    assert np.all(wat % 6 == 0)                  # (H2O)k: 6k
    assert B.workload(10, seed=4)[3][1].tolist() == mols[3][1].tolist()  # seeded


@pytest.mark.parametrize("r_near", [6.0, 30.0])
def test_mozyme_system_from_geometry_matches_the_conventional_build(r_near):
    """The sparse build (atom data, near-pair integrals, sparse H) equals the
    one cut out of the conventional system's full W and H."""
    from mopac_gpu_amd import mozyme
    nat, xyz = geometry.glycine_helix(6, 3)
    s = hamiltonian.build(nat, xyz)
    a = mozyme.MozymeSystem.from_geometry(nat, xyz, r_near=r_near, r_dipole=r_near + 1.0)
    b = mozyme.MozymeSystem(s, r_near=r_near, r_dipole=r_near + 1.0)
    assert np.array_equal(a.nijbo, b.nijbo) and a.mpack == b.mpack
    assert np.array_equal(a.w, mozyme.mozyme_integrals(s, a.nijbo))
    assert np.array_equal(a.h, b.h) and a.enuclr == b.enuclr
    assert a.s.atheat == s.atheat and a.s.nclose == s.nclose
    assert np.array_equal(a.s.pdiag, s.pdiag)
    # the nijbo layout: blocks row by row, each the size of its pair
    io = (s.nlast - s.nfirst + 1).astype(np.int64)
    off = 0
    for i in range(s.numat):
        for j in range(i + 1):
            if a.nijbo[i, j] >= 0:
                assert a.nijbo[i, j] == off and a.nijbo[j, i] == off
                off += io[i] * (io[i] + 1) // 2 if i == j else io[i] * io[j]
    assert off == a.mpack


@pytest.mark.parametrize("kind", ["helix", "water", "alkane"])
def test_lewis_lmos_are_orthonormal_and_fill_the_shells(kind):
    """The Lewis-structure start: every valence orbital used once (an
    orthogonal transformation of the AO basis), one occupied LMO per electron
    pair, and a determinant whose energy lies above the converged SCF's."""
    from mopac_gpu_amd import mozyme
    nat, xyz = {"helix": lambda: geometry.glycine_helix(3, 3),
                "water": lambda: geometry.water_cluster(4, 1),
                "alkane": lambda: geometry.alkane(5, 2)}[kind]()
    s = hamiltonian.build(nat, xyz)
    ms = mozyme.MozymeSystem.from_geometry(nat, xyz, r_near=30.0, r_dipole=31.0)
    lm = mozyme.lewis_lmos(ms)
    assert lm.nocc == s.nclose and lm.nocc + lm.nvir == s.norbs
    c = np.column_stack([mozyme.lmo_dense(lm, k, True) for k in range(lm.nocc)] +
                        [mozyme.lmo_dense(lm, k, False) for k in range(lm.nvir)])
    assert np.max(np.abs(c.T @ c - np.eye(s.norbs))) < 1e-13
    # every LMO lives on one or two atoms
    assert set(lm.ncf) <= {1, 2} and set(lm.nce) <= {2}
    p = 2.0 * c[:, :lm.nocc] @ c[:, :lm.nocc].T
    pk = np.concatenate([p[i, :i + 1] for i in range(s.norbs)])
    f = oracle.fock2(s, s.h, pk, 0.5 * pk)
    e_lewis = 2.0 * oracle.helect(s.norbs, 0.5 * pk, s.h, f)
    r = oracle.iterate(s, scfcrt=1e-6)
    assert r["converged"] and e_lewis > r["ee"]
    # with polarization from a Fock matrix the start is lower still
    fs = mozyme.sparse_from_dense(_dense(f, s.norbs), ms.nfirst, ms.nlast - ms.nfirst + 1, ms.nijbo, ms.mpack)
    lp = mozyme.lewis_lmos(ms, fs)
    c2 = np.column_stack([mozyme.lmo_dense(lp, k, True) for k in range(lp.nocc)])
    p2 = 2.0 * c2 @ c2.T
    pk2 = np.concatenate([p2[i, :i + 1] for i in range(s.norbs)])
    e2 = 2.0 * oracle.helect(s.norbs, 0.5 * pk2, s.h, oracle.fock2(s, s.h, pk2, 0.5 * pk2))
    assert r["ee"] < e2 < e_lewis


def _dense(packed, n):
    m = np.zeros((n, n))
    k = 0
    for i in range(n):
        m[i, :i + 1] = packed[k:k + i + 1]
        k += i + 1
    return m + np.tril(m, -1).T


# ---- the H, C, N, O learned guess (guess_hcno): host logic ---------------------

def _hcno_qtab():
    from mopac_gpu_amd import guess_hcno as H
    m = H.load_default()
    if m is not None:
        return m.qtab
    return np.array([0, 1], dtype=np.int64), np.array([0.1, -0.2])


@pytest.mark.parametrize("make", [lambda: geometry.glycine_helix(4, 5), lambda: geometry.water_cluster(6, 2),
                                  lambda: geometry.alkane(5, 3), lambda: geometry.water_cluster(1, 2)])
def test_hcno_descriptors_turn_with_the_molecule(make):
    """The same pairs in the same order and the same descriptors (frames,
    bonded shells and the long-range field are all taken in atom frames)
    after a rotation and a shift, including symmetric groups with equal bonds."""
    from scipy.spatial.transform import Rotation
    from mopac_gpu_amd import guess_hcno as H
    nat, xyz = make()
    qtab = _hcno_qtab()
    p1 = H.prepare(nat, xyz, qtab)
    for seed in range(3):
        Rm = Rotation.random(random_state=seed).as_matrix()
        p2 = H.prepare(nat, (xyz - xyz.mean(0)) @ Rm.T + np.array([1.0, -2.0, 0.5]), qtab)
        assert np.array_equal(p1["a"], p2["a"]) and np.array_equal(p1["b"], p2["b"])
        assert np.max(np.abs(p1["X"] - p2["X"])) < 1e-9
        assert np.array_equal(p1["typ"], p2["typ"])


def test_hcno_blocks_round_trip():
    """Blocks taken from any symmetric density in the atom frames assemble
    back to it on every covered slot (pairs within 4 A and self pairs), and
    to zero beyond."""
    from mopac_gpu_amd import guess_hcno as H
    nat, xyz = geometry.glycine_helix(3, 7)
    s = hamiltonian.build(nat, xyz)
    p = np.random.default_rng(0).standard_normal(s.mpack)
    prep = H.prepare(nat, xyz)
    back = H.assemble(s, H.blocks_from_density(s, p, prep), prep, renormalize=False)
    covered = back != 0
    assert np.max(np.abs(back[covered] - p[covered])) < 1e-12
    # the covered slots are exactly the orbital pairs of atoms within 4 A
    nf, nl = s.nfirst - 1, s.nlast
    atom = np.concatenate([np.full(nl[i] - nf[i], i) for i in range(len(nat))])
    r, c = np.tril_indices(s.norbs)
    near = np.linalg.norm(xyz[atom[r]] - xyz[atom[c]], axis=1) < H.RCUT
    assert np.array_equal(covered[r * (r + 1) // 2 + c], near)


def test_hcno_renormalized_guess_holds_the_valence_electrons():
    from mopac_gpu_amd import guess_hcno as H
    nat, xyz = geometry.glycine_helix(3, 7)
    s = hamiltonian.build(nat, xyz)
    prep = H.prepare(nat, xyz)
    y = np.abs(np.random.default_rng(1).standard_normal((len(prep["a"]), 16)))
    p = H.assemble(s, y, prep)
    d = np.arange(s.norbs) * (np.arange(s.norbs) + 1) // 2 + np.arange(s.norbs)
    assert abs(p[d].sum() - 2 * s.nclose) < 1e-9


def test_hcno_field_of_point_charges_closed_form():
    """Potential q/r and field q (r_i - r_j) / r^3 of charges beyond 4 A,
    expressed in each atom's frame; nothing from atoms within 4 A."""
    from mopac_gpu_amd import guess_hcno as H
    xyz = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [0.0, 3.0, 0.0]])
    q = np.array([0.3, -0.5, 0.7])
    R = np.tile(np.eye(3), (3, 1, 1))
    R[0] = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    f = H.field_descriptors(xyz, R, q)
    assert abs(f[0, 0] - (-0.5 / 5.0)) < 1e-14
    e0 = -0.5 * (xyz[0] - xyz[1]) / 125.0
    assert np.allclose(f[0, 1:], R[0] @ e0, atol=1e-14)
    r12 = np.linalg.norm(xyz[1] - xyz[2])
    assert abs(f[1, 0] - (0.3 / 5.0 + 0.7 / r12)) < 1e-14
    assert abs(f[2, 0] - (-0.5 / r12)) < 1e-14  # atom 0 is 3 A away: excluded


def test_hcno_charges_sum_to_the_net_charge():
    from mopac_gpu_amd import guess_hcno as H
    nat, xyz = geometry.water_cluster(2, 1)
    s = hamiltonian.build(nat, xyz)
    p = np.zeros(s.mpack)
    d = np.arange(s.norbs) * (np.arange(s.norbs) + 1) // 2 + np.arange(s.norbs)
    p[d] = 2.0 * s.nclose / s.norbs
    q = H.charges_from_density(s, p)
    assert abs(q.sum()) < 1e-12
    nbrs, _ = H.bonded_neighbours(nat, xyz)
    t = H.atom_types(nat, nbrs)
    # water: O with two bonded H; H with one bonded O
    assert len(set(t[nat == 8])) == 1 and len(set(t[nat == 1])) == 1
    assert t[nat == 8][0] == 3 + 4 * 2 and t[nat == 1][0] == 0 + 4 * 125
