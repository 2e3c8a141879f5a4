"""GPU parity: every hot-path routine on the HIP path against the CPU oracle
(oracle/oracle_scf.c), plus the full SCF against the oracle and the
reference's own golden outputs."""
import ctypes
import json
import os

import numpy as np
import pytest

from mopac_gpu_amd import geometry, hamiltonian
from mopac_gpu_amd import lib as L
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.json")))


def _system(name):
    g = GOLD[name]
    return hamiltonian.build(np.array(g["nat"]), np.array(g["xyz"]))


def _water(k, seed):
    nat, xyz = geometry.water_cluster(k, seed)
    return hamiltonian.build(nat, xyz)


def _rand_density(sysm, seed):
    rng = np.random.default_rng(seed)
    n = sysm.norbs
    a = rng.standard_normal((n, n))
    q, _ = np.linalg.qr(a)
    return oracle.density(np.asfortranarray(q), sysm.nclose)


def _pk_diag(n):
    return np.array([i * (i + 1) // 2 + i for i in range(n)])


@pytest.mark.parametrize("which", ["c2h4", "water4"])
def test_fock2_matches_oracle(gpu_ctx, which):
    s = _system("c2h4_1scf") if which == "c2h4" else _water(4, 11)
    gpu_ctx.upload(s)
    p = _rand_density(s, 3)
    pa = 0.5 * p
    pa[_pk_diag(s.norbs)] *= 0.97  # alpha density deliberately not P/2
    f_gpu = gpu_ctx.fock2(s.h, p, pa)
    f_ref = oracle.fock2(s, s.h, p, pa)
    assert np.max(np.abs(f_gpu - f_ref)) < 1e-11 * max(1.0, np.max(np.abs(f_ref)))


@pytest.mark.parametrize("nw", [8, 64, 256])
def test_density_and_energy_match_oracle(gpu_ctx, nw):
    """(H2O)64 (n = 384, 256 occupied) takes the large-tile packed SYRK, (H2O)256
    (n = 1536, 1024 occupied) that SYRK with K split in four (packed parts)."""
    s = _water(nw, 5)
    gpu_ctx.upload(s)
    rng = np.random.default_rng(1)
    q, _ = np.linalg.qr(rng.standard_normal((s.norbs, s.norbs)))
    c = np.asfortranarray(q)
    p_gpu = gpu_ctx.density(c, s.nclose, 2.0)
    p_ref = oracle.density(c, s.nclose, 2.0)
    assert np.max(np.abs(p_gpu - p_ref)) < 1e-13 * np.sqrt(s.nclose)
    # an occupied count that is not a multiple of the k step takes the 64 x 64 tiles
    p_odd = gpu_ctx.density(c, s.nclose - 3, 2.0)
    assert np.max(np.abs(p_odd - oracle.density(c, s.nclose - 3, 2.0))) < 1e-13 * np.sqrt(s.nclose)
    f = oracle.fock2(s, s.h, p_ref, 0.5 * p_ref)
    e_gpu = gpu_ctx.helect(0.5 * p_ref, s.h, f)
    e_ref = oracle.helect(s.norbs, 0.5 * p_ref, s.h, f)
    assert abs(e_gpu - e_ref) < 1e-10 * max(1.0, abs(e_ref))


@pytest.mark.parametrize("schedule", [1, 0])
@pytest.mark.parametrize("k", [8, 32, 64])
def test_eigen_matches_lapack(gpu_ctx, k, schedule):
    s = _water(k, 7)
    gpu_ctx.upload(s)
    p = _rand_density(s, 9)
    f = oracle.fock2(s, s.h, p, 0.5 * p)
    f1, f2 = f.copy(), f.copy()
    print(f"eigen (H2O){k} schedule {schedule}", flush=True)
    lib = L.load()
    prev = lib.mhip_set_jacobi_schedule(schedule)
    try:
        c_gpu, e_gpu = gpu_ctx.eigenvectors(f1)
    finally:
        lib.mhip_set_jacobi_schedule(prev)
    c_ref, e_ref = oracle.eigenvectors(f2)
    # the reference perturbs its argument in place; both must agree
    assert np.array_equal(f1, f2)
    assert np.max(np.abs(e_gpu - e_ref)) < 1e-9
    n = s.norbs
    fm = np.zeros((n, n))
    iu = np.tril_indices(n)
    fm[iu] = f1
    fm = fm + fm.T - np.diag(np.diag(fm))
    assert np.max(np.abs(fm @ c_gpu - c_gpu * e_gpu)) < 1e-9
    assert np.max(np.abs(c_gpu.T @ c_gpu - np.eye(n))) < 1e-12
    # same occupied subspace => same density
    no = s.nclose
    assert np.max(np.abs(c_gpu[:, :no] @ c_gpu[:, :no].T - c_ref[:, :no] @ c_ref[:, :no].T)) < 1e-9


def test_scf_co2_reproduces_reference_outputs(gpu_ctx):
    g = GOLD["co2_1scf"]
    s = _system("co2_1scf")
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(pulay=False, scfcrt=1e-4)
    o = oracle.iterate(s, pulay=False, scfcrt=1e-4)
    assert r["converged"] and o["converged"]
    assert r["niter"] == o["niter"]
    assert abs(r["escf"] - o["escf"]) < 1e-6
    assert np.max(np.abs(r["eig"] - np.array(g["eigenvalues"]))) < 6e-6
    assert np.max(np.abs(r["p"] - o["p"])) < 1e-6


def test_scf_h2o_pulay_reproduces_reference_charges(gpu_ctx):
    g = GOLD["h2o_api"]
    s = _system("h2o_api")
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(pulay=True, scfcrt=1e-4)
    o = oracle.iterate(s, pulay=True, scfcrt=1e-4)
    assert r["converged"] and r["niter"] == o["niter"]
    p = r["p"]
    d = _pk_diag(s.norbs)
    q = [1 - p[d[0]], 1 - p[d[1]], 6 - p[d[2:6]].sum()]
    assert np.max(np.abs(np.array(q) - np.array(g["charges"]))) < 6e-6
    assert abs(r["escf"] - o["escf"]) < 1e-6


@pytest.mark.parametrize("k", [4, 8])
def test_pulay_sequence_wraps_history(gpu_ctx, k):
    """14 consecutive Pulay steps on the device (more than the 10 stored Fock
    matrices, so the circular history wraps and the full 11 x 11 DIIS system
    is solved) against the oracle's pulay, which equals the reference's
    Fortran over the same kind of sequence; random densities, Fock matrices
    built from them."""
    s = _water(k, 3)
    gpu_ctx.upload(s)
    ncall = 14
    fs = np.stack([oracle.fock2(s, s.h, _rand_density(s, 30 + i), 0.5 * _rand_density(s, 30 + i))
                   for i in range(ncall)])
    ps = np.stack([0.5 * _rand_density(s, 50 + i) for i in range(ncall)])
    f_o, pl_o = oracle.pulay_seq(s.norbs, fs, ps)
    f_g = np.ascontiguousarray(fs.copy())
    pl_g = np.zeros(ncall)
    L.check(L.load().mhip_pulay_sequence(gpu_ctx.h, ncall, L.ptr(f_g), L.ptr(np.ascontiguousarray(ps)),
                                         L.ptr(pl_g)))
    assert np.max(np.abs(pl_g - pl_o)) <= 1e-12 * np.max(np.abs(pl_o))
    assert np.max(np.abs(f_g - f_o)) <= 1e-10 * np.max(np.abs(f_o))


def test_scf_pulay_matches_oracle(gpu_ctx):
    """PULAY on (H2O)32 at SCFCRT = 1e-8, the device eigensolver at its
    tightest tolerance so the trajectory follows the oracle's (LAPACK)."""
    nat, xyz = geometry.water_cluster(32, 1, shape=(4, 4, 2))
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    lib = L.load()
    lib.mhip_set_tuning(gpu_ctx.h, 1, -1, 1.0e-13)
    try:
        r = gpu_ctx.iter(pulay=True, scfcrt=1e-8)
    finally:
        lib.mhip_set_tuning(gpu_ctx.h, 1, -1, 1.0e-10)
    o = oracle.iterate(s, pulay=True, scfcrt=1e-8)
    assert r["converged"] and o["converged"]
    assert r["niter"] == o["niter"]
    assert abs(r["escf"] - o["escf"]) < 1e-6
    assert np.max(np.abs(r["p"] - o["p"])) < 1e-7


def test_scf_water32_matches_oracle(gpu_ctx):
    """Config 1: (H2O)32, 192 basis functions, SCFCRT=1e-6."""
    nat, xyz = geometry.water_cluster(32, 1, shape=(4, 4, 2))
    s = hamiltonian.build(nat, xyz)
    assert s.norbs == 192 and s.nclose == 128
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(pulay=False, scfcrt=1e-6)
    o = oracle.iterate(s, pulay=False, scfcrt=1e-6)
    assert r["converged"] and o["converged"]
    assert abs(r["escf"] - o["escf"]) < 1e-3
    assert np.max(np.abs(r["p"] - o["p"])) < 1e-6
    assert np.max(np.abs(r["eig"] - o["eig"])) < 1e-6
    assert r["niter"] == o["niter"]


def test_pseudo_diag_matches_oracle(gpu_ctx):
    s = _water(8, 21)
    gpu_ctx.upload(s)
    o = oracle.iterate(s, pulay=False, scfcrt=1e-3)
    p = o["p"] * 1.0
    rng = np.random.default_rng(4)
    p += 1e-3 * rng.standard_normal(p.shape)
    f = oracle.fock2(s, s.h, p, 0.5 * p)
    v_gpu, nr_gpu = gpu_ctx.diag(f, o["c"], s.nclose, o["eig"])
    v_ref, nr_ref = oracle.diag(f, o["c"], s.nclose, o["eig"])
    assert nr_gpu == nr_ref and nr_ref > 0
    assert np.max(np.abs(v_gpu - v_ref)) < 1e-12


def test_reference_boundary_dropins(gpu_ctx):
    lib = L.load()
    rng = np.random.default_rng(2)
    n = 96
    a = rng.standard_normal((n, n))
    a = np.asfortranarray(a + a.T)
    w = np.zeros(n)
    info = ctypes.c_int()
    aa = a.copy(order="F")
    lib.mhip_dsyevd(n, L.ptr(aa), n, L.ptr(w), ctypes.byref(info))
    assert info.value == 0
    assert np.max(np.abs(w - np.linalg.eigvalsh(a))) < 1e-10
    assert np.max(np.abs(a @ aa - aa * w)) < 1e-9
    k = 40
    x = np.asfortranarray(rng.standard_normal((n, k)))
    cc = np.zeros((n, n), order="F")
    lib.mhip_call_syrk(b"U", b"N", n, k, 2.0, L.ptr(x), n, 0.0, L.ptr(cc), n)
    ref = 2.0 * x @ x.T
    assert np.max(np.abs(np.triu(cc) - np.triu(ref))) < 1e-12
    b = np.asfortranarray(rng.standard_normal((k, n)))
    cg = np.asfortranarray(rng.standard_normal((n, n)))
    ref = 1.5 * x @ b + 0.5 * cg
    lib.mhip_call_gemm(b"N", b"N", n, n, k, 1.5, L.ptr(x), n, L.ptr(b), k, 0.5, L.ptr(cg), n)
    assert np.max(np.abs(cg - ref)) < 1e-12


@pytest.mark.parametrize("tra,n,k", [("T", 64, 48), ("N", 640, 256), ("T", 1024, 3072), ("T", 2048, 3072),
                                     ("N", 1024, 1024)])
def test_symmetric_result_gemm_matches_numpy(gpu_ctx, tra, n, k):
    """call_syrk's product op(A) op(A)^T on the lower-triangle tiles: the
    64- and 128-wide tiles, with and without the K split (the Bjorck Gram
    shapes of (H2O)512 are n = 1024 and 2048 at k = 3072); both triangles
    against numpy, and the full product exactly symmetric."""
    lib = L.load()
    rng = np.random.default_rng(n + k)
    x = np.asfortranarray(rng.standard_normal((k, n) if tra == "T" else (n, k)))
    ref = 0.5 * (x.T @ x if tra == "T" else x @ x.T)
    tol = 1e-12 * np.sqrt(k) * np.max(np.abs(ref))
    for uplo in (b"U", b"L"):
        cc = np.zeros((n, n), order="F")
        lib.mhip_call_syrk(uplo, tra.encode(), n, k, 0.5, L.ptr(x), x.shape[0], 0.0, L.ptr(cc), n)
        tri = np.triu if uplo == b"U" else np.tril
        assert np.max(np.abs(tri(cc) - tri(ref))) < tol
        # the other triangle is left as it was
        other = np.tril(cc, -1) if uplo == b"U" else np.triu(cc, 1)
        assert not np.any(other)
    # the mirrored upper triangle equals the lower one bit for bit
    cu = np.zeros((n, n), order="F")
    cl = np.zeros((n, n), order="F")
    lib.mhip_call_syrk(b"U", tra.encode(), n, k, 0.5, L.ptr(x), x.shape[0], 0.0, L.ptr(cu), n)
    lib.mhip_call_syrk(b"L", tra.encode(), n, k, 0.5, L.ptr(x), x.shape[0], 0.0, L.ptr(cl), n)
    assert np.array_equal(np.triu(cu), np.tril(cl).T)


def test_mfma_tiles_match_vector_tiles(gpu_ctx):
    """Checks the FP64 MFMA lane layout: both kernel families against numpy."""
    lib = L.load()
    rng = np.random.default_rng(8)
    m, n, k = 150, 130, 70
    a = np.asfortranarray(rng.standard_normal((k, m)))
    b = np.asfortranarray(rng.standard_normal((k, n)))
    ref = a.T @ b
    outs = []
    for variant in (0, 1):
        lib.mhip_set_tuning(gpu_ctx.h, variant, -1, -1.0)
        c = np.zeros((m, n), order="F")
        lib.mhip_call_gemm(b"T", b"N", m, n, k, 1.0, L.ptr(a), k, L.ptr(b), k, 0.0, L.ptr(c), m)
        outs.append(c)
    lib.mhip_set_tuning(gpu_ctx.h, 1, -1, -1.0)
    assert np.max(np.abs(outs[0] - ref)) < 1e-12
    assert np.max(np.abs(outs[1] - ref)) < 1e-12


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(256, 256, 64), (384, 192, 272), (64, 64, 16), (256, 128, 3072),
                                   (1024, 1024, 1536), (2048, 1024, 3072)])
def test_large_tile_gemm_matches_numpy(gpu_ctx, ta, tb, m, n, k):
    """Shapes that divide the 128x128, 128x64 or 64x64 tiles take the
    double-buffered large-tile kernel; alpha and beta on a random C.  The
    launcher splits K for CU balance: (256, 128, 3072) takes 4 slices,
    (1024, 1024, 1536) 2 and (2048, 1024, 3072) 3, the first slice adding
    beta C and the others summed in after it."""
    lib = L.load()
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    a = np.asfortranarray(rng.standard_normal((k, m) if ta == "T" else (m, k)))
    b = np.asfortranarray(rng.standard_normal((n, k) if tb == "T" else (k, n)))
    c = np.asfortranarray(rng.standard_normal((m, n)))
    opa = a.T if ta == "T" else a
    opb = b.T if tb == "T" else b
    ref = 0.75 * opa @ opb - 0.5 * c
    lib.mhip_call_gemm(ta.encode(), tb.encode(), m, n, k, 0.75, L.ptr(a), a.shape[0], L.ptr(b), b.shape[0],
                       -0.5, L.ptr(c), m)
    assert np.max(np.abs(c - ref)) < 1e-12 * max(1.0, np.sqrt(k))
