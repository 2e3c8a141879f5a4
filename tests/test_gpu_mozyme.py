"""GPU parity of the MOZYME kernels (density_for_MOZYME, diagg1, diagg2)
against MOPAC's own Fortran routines compiled from the reference sources
(oracle/_ref, oracle/ref.py), on localized MOs of a converged poly-glycine
helix, over a sequence of calls that exercises every branch of the saved
state (new lists, reused lists, occupied energies, tiny/biglim, growth)."""
import numpy as np
import pytest

from mopac_gpu_amd import geometry, hamiltonian, mozyme
from oracle import oracle
from oracle import ref as REF

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not REF.available(), reason="oracle/_ref not built (needs the reference sources)")]


def _unpack(p, n):
    m = np.zeros((n, n))
    m[np.tril_indices(n)] = p
    return m + m.T - np.diag(np.diag(m))


@pytest.fixture(scope="module")
def helix_lmos():
    nat, xyz = geometry.glycine_helix(3, 3)
    s = hamiltonian.build(nat, xyz)
    o = oracle.iterate(s, scfcrt=1e-6)
    n = s.norbs
    return mozyme.from_scf(s, o["c"], _unpack(o["f"], n), _unpack(o["p"], n), s.nclose, keep=1e-3, rcut=6.0)


def _same_lmos(a, b):
    assert np.array_equal(a.ncf, b.ncf) and np.array_equal(a.nce, b.nce)
    for k in range(a.nocc):
        sl = slice(a.nncf[k], a.nncf[k] + a.ncf[k])
        assert np.array_equal(a.icocc[sl], b.icocc[sl])
    for k in range(a.nvir):
        sl = slice(a.nnce[k], a.nnce[k] + a.nce[k])
        assert np.array_equal(a.icvir[sl], b.icvir[sl])
    assert np.max(np.abs(a.cocc - b.cocc)) < 1e-12
    assert np.max(np.abs(a.cvir - b.cvir)) < 1e-12


def test_density_for_mozyme_matches_reference(gpu_ctx, helix_lmos):
    L = helix_lmos
    print(f"helix LMOs: {L.numat} atoms, {L.nocc} occupied, {L.nvir} virtual", flush=True)
    R = REF.Mozyme(L)
    G = mozyme.GpuMozyme(gpu_ctx, L)
    for mode in (0, -1, 1):
        part = None if mode == 0 else L.p
        pr = R.density(mode, part)
        pg = G.density(mode, part)
        assert np.max(np.abs(pg - pr)) < 1e-13 * max(1.0, np.max(np.abs(pr)))


@pytest.fixture(scope="module")
def lewis_helix_lmos():
    """Lewis-structure LMOs of a 12-residue helix (87 atoms, 6 A stored pairs),
    F and P of the diagonal starting density from the oracle: a longer rotation
    list, with many rotations that share no LMO (the level-parallel diagg2),
    and little room per LMO, so growth is rejected and restored."""
    nat, xyz = geometry.glycine_helix(12, 3)
    s = hamiltonian.build(nat, xyz)
    ms = mozyme.MozymeSystem(s, r_near=6.0, r_dipole=7.0)
    p0 = ms.diagonal_density()
    f0 = oracle.mozyme_fock2z(s, ms.nijbo, ms.mpack, ms.h, ms.am, ms.ad, ms.dd, p0, feather=True)
    L = mozyme.lewis_lmos(ms, f0, pad_atoms=2)
    L.fao, L.p = f0, p0
    fd = mozyme.dense_from_sparse(f0, ms.nfirst, ms.nlast - ms.nfirst + 1, ms.nijbo, ms.norbs)
    L.eigs = np.array([mozyme._energy(L, k, True, fd) for k in range(L.nocc)])
    L.eigv = np.array([mozyme._energy(L, k, False, fd) for k in range(L.nvir)])
    return L


@pytest.fixture(params=[(1, 1, 1024, 1024), (1, 1 << 30, 1024, 1024), (1, 0, 1024, 1024), (1, 1, 0, 1024),
                        (1, 1, 1024, 0), (0, 1, 1024, 1024)],
                ids=["lds_hybrid", "lds_one_workgroup", "lds_across_chip", "lds_lists_in_global",
                     "lds_coefficients_in_global", "global_per_level"])
def diagg2_path(request):
    """Every way of running diagg2's levels: the LDS kernels with the default
    split, with every level in the one workgroup, with every level across the
    chip, with the atom lists read from global memory, with the coefficients
    read from global memory (the lists in LDS, an undo log for rejections),
    and the global-map kernel one launch per level."""
    from mopac_gpu_amd import lib as L_
    lib = L_.load()
    was = lib.mhip_mozyme_diagg2_one_workgroup(request.param[0])
    was_w = lib.mhip_mozyme_diagg2_wide_rounds(request.param[1])
    was_c = lib.mhip_mozyme_diagg2_list_cap(request.param[2])
    was_p = lib.mhip_mozyme_diagg2_coef_cap(request.param[3])
    yield request.param
    lib.mhip_mozyme_diagg2_one_workgroup(was)
    lib.mhip_mozyme_diagg2_wide_rounds(was_w)
    lib.mhip_mozyme_diagg2_list_cap(was_c)
    lib.mhip_mozyme_diagg2_coef_cap(was_p)


def test_diagg_sequence_on_lewis_lmos_matches_reference(gpu_ctx, lewis_helix_lmos, diagg2_path):
    _diagg_sequence(gpu_ctx, lewis_helix_lmos, None, 8, diagg2_path)


@pytest.mark.parametrize("fmo_cap", [None, 60])
def test_diagg1_diagg2_sequence_matches_reference(gpu_ctx, helix_lmos, fmo_cap, diagg2_path):
    _diagg_sequence(gpu_ctx, helix_lmos, fmo_cap, 10, diagg2_path)


def _diagg_sequence(gpu_ctx, L, fmo_cap, ncalls, path=1):
    R = REF.Mozyme(L, fmo_dim=fmo_cap)
    G = mozyme.GpuMozyme(gpu_ctx, L, fmo_dim=fmo_cap)
    for idiagg in range(1, ncalls + 1):
        print(f"idiagg {idiagg} cap {fmo_cap} ({L.numat} atoms) diagg2 path {path}", flush=True)
        r = R.diagg1(L.fao, idiagg)
        g = G.diagg1(L.fao, L.p, idiagg)
        assert g["nij"] == r["nij"]
        assert np.array_equal(g["ifmo"], r["ifmo"])
        assert np.max(np.abs(g["fmo"] - r["fmo"]), initial=0.0) < 1e-12
        assert np.max(np.abs(g["eigv"] - r["eigv"])) < 1e-12
        assert g["tiny"] == pytest.approx(r["tiny"], rel=1e-12, abs=1e-300)
        assert g["sumt"] == pytest.approx(r["sumt"], rel=1e-12, abs=1e-300)
        sr = R.diagg2(r["nij"], idiagg)
        sg = G.diagg2(idiagg)
        assert sg == pytest.approx(sr, rel=1e-12, abs=1e-300)
        a, b = R.get(), G.get()
        assert np.max(np.abs(a.eigs - b.eigs)) < 1e-12
        _same_lmos(a, b)


def _idiag(L):
    out = []
    for a in range(L.numat):
        for l1 in range(L.iorbs[a]):
            out.append(L.nijbo[a, a] + l1 * (l1 + 1) // 2 + l1 + 1)
    return np.array(out, dtype=np.int32)


def test_helecz_matches_reference(gpu_ctx, helix_lmos):
    import ctypes
    from mopac_gpu_amd import lib as L_
    L = helix_lmos
    R = REF.Mozyme(L)
    rng = np.random.default_rng(3)
    h = L.fao - rng.uniform(0.0, 1.0, L.mpack)
    e_ref = R.helecz(L.p, h, L.fao)
    e = ctypes.c_double()
    nij = np.asfortranarray(L.nijbo.astype(np.int32))
    L_.check(L_.load().mhip_mozyme_helecz(gpu_ctx.h, L.numat, L_.ptr(L.nfirst.astype(np.int32)),
                                          L_.ptr(L.nlast.astype(np.int32)), L_.ptr(nij), L.mpack, L_.ptr(L.p),
                                          L_.ptr(h), L_.ptr(L.fao), ctypes.byref(e)))
    assert e.value == pytest.approx(e_ref, rel=1e-13)


@pytest.mark.parametrize("niter,three", [(2, 1), (3, 1), (6, 1), (7, 0)])
def test_cnvgz_matches_reference(gpu_ctx, helix_lmos, niter, three):
    import ctypes
    from mopac_gpu_amd import lib as L_
    L = helix_lmos
    R = REF.Mozyme(L)
    rng = np.random.default_rng(niter)
    idiag = _idiag(L)
    base = [L.p + 0.2 * rng.standard_normal(L.mpack), L.p.copy()] + \
           [L.p[idiag - 1] + 0.1 * rng.standard_normal(L.norbs) for _ in range(3)]
    ref_arrays = [b.copy() for b in base]
    gpu_arrays = [np.ascontiguousarray(b.copy()) for b in base]
    pm_ref = R.cnvgz(*ref_arrays, niter, idiag, three_point=bool(three))
    pm = ctypes.c_double()
    L_.check(L_.load().mhip_mozyme_cnvgz(gpu_ctx.h, L.norbs, L.mpack, *[L_.ptr(a) for a in gpu_arrays], niter,
                                         L_.ptr(idiag), three, ctypes.byref(pm)))
    assert pm.value == pm_ref
    for a, b in zip(gpu_arrays, ref_arrays):
        assert np.max(np.abs(a - b)) < 1e-13


@pytest.mark.parametrize("r_near,r_dipole,feather", [(6.0, 7.0, False), (4.0, 6.0, False), (30.0, 31.0, False),
                                                     (4.0, 6.0, True), (6.0, 8.0, True)])
def test_fock2z_matches_reference(gpu_ctx, r_near, r_dipole, feather):
    """buildf/fock2z on the device (stored pairs, dipole shell, point charges)
    against the reference's own fz2n, SEMIDIRECT, and against the oracle."""
    nat, xyz = geometry.glycine_helix(6, 3)
    s = hamiltonian.build(nat, xyz)
    ms = mozyme.MozymeSystem(s, r_near=r_near, r_dipole=r_dipole, feather=feather)
    rng = np.random.default_rng(1)
    pd = rng.standard_normal((s.norbs, s.norbs)) * 0.1
    pd = pd + pd.T + np.eye(s.norbs)
    p = mozyme.sparse_from_dense(pd, ms.nfirst, ms.nlast - ms.nfirst + 1, ms.nijbo, ms.mpack)
    print(f"fock2z helix {ms.numat} atoms, near {r_near}, dipole {r_dipole}", flush=True)
    ms.upload(gpu_ctx)
    f = ms.fock2z(gpu_ctx, p)
    f_ref = REF.mozyme_fock2z(ms, p, feather=feather)
    scale = np.max(np.abs(f_ref))
    assert np.max(np.abs(f - f_ref)) < 1e-11 * scale
    f_o = oracle.mozyme_fock2z(s, ms.nijbo, ms.mpack, ms.h, ms.am, ms.ad, ms.dd, p, feather=feather)
    assert np.max(np.abs(f - f_o)) < 1e-11 * scale


@pytest.mark.parametrize("r_near,r_dipole", [(30.0, 31.0), (8.0, 9.0)])
def test_mozyme_scf_reaches_the_conventional_energy(gpu_ctx, r_near, r_dipole):
    """iter_for_MOZYME on the device from LMOs of an unconverged start (two
    conventional iterations), with every atom pair stored or with point
    charges and dipoles beyond 8 A: at MOZYME's default criterion (SELCON
    0.01 kcal/mol) the heat of formation is the conventional SCF's within
    0.05 kcal/mol."""
    nat, xyz = geometry.glycine_helix(6, 3)
    s = hamiltonian.build(nat, xyz)
    n = s.norbs
    gpu_ctx.upload(s)
    conv = gpu_ctx.iter(scfcrt=1e-6)
    r2 = gpu_ctx.iter(scfcrt=1e-6, itrmax=2, want=("p", "f", "c", "eig"))
    fd, pd = _unpack(r2["f"], n), _unpack(r2["p"], n)
    lmos = mozyme.from_scf(s, r2["c"], fd, pd, s.nclose, keep=1e-8, rcut=r_near, r_dipole=r_dipole)
    assert (lmos.nijbo >= 0).all() == (r_near > 20)
    ms = mozyme.MozymeSystem(s, nijbo=lmos.nijbo, mpack=lmos.mpack)
    ms.upload(gpu_ctx)
    print(f"MOZYME SCF helix {s.numat} atoms from a two-iteration start", flush=True)
    r = ms.scf(gpu_ctx, lmos)
    print(f"escf {r['escf']:.6f} conventional {conv['escf']:.6f} niter {r['niter']}", flush=True)
    assert r["converged"] and 3 < r["niter"] < 100
    assert abs(r["escf"] - conv["escf"]) < 0.05
    # the density left by the loop carries 2 electrons per occupied LMO
    d = np.array([ms.nijbo[a, a] + (k * (k + 1)) // 2 + k for a in range(s.numat)
                  for k in range(int(s.nlast[a] - s.nfirst[a] + 1))])
    assert abs(np.sum(r["p"][d]) - 2 * s.nclose) < 1e-3 * s.nclose


def test_lewis_start_reaches_the_conventional_energy(gpu_ctx):
    """The makvec-style start: Lewis-structure LMOs polarized by the Fock
    matrix of the diagonal density (built by the device fock2z), on a system
    built straight from the geometry.  With every pair stored the MOZYME SCF
    ends within 0.05 kcal/mol of the conventional SCF at MOZYME's default
    criterion, as the start from localized conventional orbitals does."""
    nat, xyz = geometry.glycine_helix(6, 3)
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    conv = gpu_ctx.iter(scfcrt=1e-6)
    ms = mozyme.MozymeSystem.from_geometry(nat, xyz, r_near=30.0, r_dipole=31.0)
    ms.upload(gpu_ctx)
    f0 = ms.fock2z(gpu_ctx, ms.diagonal_density())
    lmos = mozyme.lewis_lmos(ms, f0)
    print(f"MOZYME SCF helix {ms.numat} atoms from Lewis LMOs", flush=True)
    r = ms.scf(gpu_ctx, lmos)
    print(f"escf {r['escf']:.6f} conventional {conv['escf']:.6f} niter {r['niter']}", flush=True)
    assert r["converged"] and 3 < r["niter"] < 60
    assert abs(r["escf"] - conv["escf"]) < 0.05


@pytest.mark.parametrize("nres,scfcrt", [(1, 1e-2), (12, 1e-2), (30, 1e-3), (400, 1e-3)])
def test_mozyme_loop_follows_the_reference_routines(gpu_ctx, nres, scfcrt):
    """The whole MOZYME SCF on the device against iter_for_MOZYME's loop
    driven over the reference's own routines (tidy, check, eimp, diagg1,
    diagg2, density_for_MOZYME, cnvgz, fz2n, helecz; oracle/ref.py
    mozyme_scf) from the same Lewis start at MOZYME's default cutoffs: Gly1
    (under 30 atoms, every pair stored), Gly12, Gly30 and the 400-residue
    helix of the benchmark (stored, dipole and point-charge pairs; about 12 s
    of the reference routines on one host core).  The heat after every
    iteration agrees to 1e-6 kcal/mol, or 1e-10 of the heat where that is
    larger (the 400-residue helix's heats reach 2.8e4 kcal/mol, and the
    energy sums over its 882 259 stored elements round in another order
    than the reference's sequential loops: 1.6e-6 kcal/mol at most, already
    6e-7 at the first iteration), and both stop at the same iteration."""
    nat, xyz = geometry.glycine_helix(nres, 3)
    ms = mozyme.MozymeSystem.from_geometry(nat, xyz)
    ms.upload(gpu_ctx)
    f0 = ms.fock2z(gpu_ctx, ms.diagonal_density())
    lmos = mozyme.lewis_lmos(ms, f0)
    print(f"MOZYME loop Gly{nres}: {ms.numat} atoms, SCFCRT {scfcrt}", flush=True)
    r = ms.scf(gpu_ctx, lmos, scfcrt=scfcrt, heats_cap=500)
    ref = REF.mozyme_scf(ms, lmos, scfcrt=scfcrt)
    print(f"gpu {r['escf']:.6f} ({r['niter']} it), reference routines {ref['escf']:.6f} ({ref['niter']} it)",
          flush=True)
    assert r["converged"] and ref["converged"]
    assert r["niter"] == ref["niter"]
    tol = max(1e-6, 1e-10 * np.max(np.abs(ref["heats"])))
    assert np.max(np.abs(r["heats"] - ref["heats"])) < tol


@pytest.mark.parametrize("nres", [12, 30])
def test_order_free_sweep_converges_to_the_reference_order_heat(gpu_ctx, nres):
    """The order-free form (diagg2's rotations in LMO-disjoint rounds) ends at
    the reference-order form's heat within the metric's 1e-3 kcal/mol, at
    SCFCRT 1e-3, with far fewer dependent rounds than the reference order's
    levels."""
    import ctypes
    from mopac_gpu_amd import lib as L_
    lib = L_.load()
    nat, xyz = geometry.glycine_helix(nres, 3)
    ms = mozyme.MozymeSystem.from_geometry(nat, xyz)
    ms.upload(gpu_ctx)
    f0 = ms.fock2z(gpu_ctx, ms.diagonal_density())
    lmos = mozyme.lewis_lmos(ms, f0)
    out = {}
    for of in (False, True):
        sv = [ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double(),
              ctypes.c_longlong()]
        lib.mhip_mozyme_stats(1, *[ctypes.byref(v) for v in sv])
        r = ms.scf(gpu_ctx, lmos, scfcrt=1e-3, order_free=of)
        lib.mhip_mozyme_stats(0, *[ctypes.byref(v) for v in sv])
        out[of] = (r, sv[2].value)
        print(f"Gly{nres} order_free={of}: escf {r['escf']:.6f}, {r['niter']} iterations, levels {sv[2].value}",
              flush=True)
    (r0, lev0), (r1, lev1) = out[False], out[True]
    assert r0["converged"] and r1["converged"]
    assert abs(r1["escf"] - r0["escf"]) < 1e-3
    assert abs(r1["niter"] - r0["niter"]) <= 5
    assert lev1 < lev0


def _lmo_extents(nc, nn, no, ic, c, iorbs):
    """Each LMO's atoms and coefficients (the parts tidy defines)."""
    atoms, coefs = [], []
    for k in range(len(nc)):
        a = ic[nn[k]:nn[k] + nc[k]]
        atoms.append(a.copy())
        coefs.append(c[no[k]:no[k] + int(sum(iorbs[x - 1] for x in a))].copy())
    return atoms, coefs


@pytest.fixture(scope="module")
def tidy_lmos(gpu_ctx):
    """LMOs of a Gly12 helix at MOZYME's cutoffs after four iterations of the
    device MOZYME SCF from the Lewis start (grown LMOs with small tails, in
    the reference's capacities), every LMO's norm moved off 1 by ~1e-5 (so
    check has something to restore while staying below the reference's 0.1
    stop)."""
    nat, xyz = geometry.glycine_helix(12, 3)
    ms = mozyme.MozymeSystem.from_geometry(nat, xyz)
    ms.upload(gpu_ctx)
    f0 = ms.fock2z(gpu_ctx, ms.diagonal_density())
    L = ms.scf(gpu_ctx, mozyme.lewis_lmos(ms, f0), itrmax=4)["lmos"]
    rng = np.random.default_rng(5)
    iorbs = L.iorbs
    for nc, nn, no, ic, c in ((L.ncf, L.nncf, L.ncocc, L.icocc, L.cocc), (L.nce, L.nnce, L.ncvir, L.icvir, L.cvir)):
        for k in range(len(nc)):
            n = int(sum(iorbs[a - 1] for a in ic[nn[k]:nn[k] + nc[k]]))
            c[no[k]:no[k] + n] *= 1.0 + 1e-5 * rng.standard_normal()
    L.eigs, L.eigv = np.zeros(L.nocc), np.zeros(L.nvir)
    return ms, L


@pytest.mark.parametrize("occupied", [1, 0])
def test_tidy_and_check_match_reference(gpu_ctx, tidy_lmos, occupied):
    """mhip_mozyme_tidy and mhip_mozyme_check on the device against the
    reference's tidy.F90 and check.F90 (oracle/_ref) on the same LMO set:
    identical counts, offsets, atom lists and coefficients after tidy; after
    check the same normalization error and coefficients to 1e-15."""
    ms, L = tidy_lmos
    iorbs = L.iorbs
    if occupied:
        args = (L.ncf, L.nncf, L.ncocc, L.icocc, L.cocc)
    else:
        args = (L.nce, L.nnce, L.ncvir, L.icvir, L.cvir)
    thresh = 1e-6  # tidy's THRESH: drops the grown LMOs' small tails
    nc, nn, no, ic, c, ln, mn, err = REF.mozyme_tidy(L.numat, L.norbs, iorbs, *args, thresh,
                                                       mode=1 if occupied else 2)
    assert err == 0
    G = mozyme.GpuMozyme(gpu_ctx, L)
    G.tidy(occupied, thresh)
    g = G.get()
    gnc, gnn, gno, gic, gc = ((g.ncf, g.nncf, g.ncocc, g.icocc, g.cocc) if occupied else
                              (g.nce, g.nnce, g.ncvir, g.icvir, g.cvir))
    print(f"tidy occupied={occupied}: {int(np.sum(args[0]) - np.sum(nc))} atoms dropped", flush=True)
    assert np.sum(nc) < np.sum(args[0])
    assert np.array_equal(gnc, nc) and np.array_equal(gnn, nn) and np.array_equal(gno, no)
    ra, rc = _lmo_extents(nc, nn, no, ic, c, iorbs)
    ga, gcf = _lmo_extents(gnc, gnn, gno, gic, gc, iorbs)
    for k in range(len(nc)):
        assert np.array_equal(ra[k], ga[k])
        assert np.array_equal(rc[k], gcf[k])
    # check on the tidied set
    cr, ws = REF.mozyme_check(L.numat, L.norbs, iorbs, nc, nn, no, ic, c)
    e = G.check(occupied)
    assert e == pytest.approx(np.sum(np.abs(1.0 - ws)), rel=1e-12)
    g = G.get()
    gc = g.cocc if occupied else g.cvir
    _, rc2 = _lmo_extents(nc, nn, no, ic, cr, iorbs)
    _, gc2 = _lmo_extents(nc, nn, no, ic, gc, iorbs)
    for k in range(len(nc)):
        assert np.max(np.abs(rc2[k] - gc2[k]), initial=0.0) < 1e-15


def test_eimp_matches_reference(gpu_ctx, tidy_lmos):
    from mopac_gpu_amd import lib as L_
    ms, L = tidy_lmos
    rng = np.random.default_rng(9)
    f = rng.standard_normal(ms.mpack)
    p = rng.standard_normal(ms.mpack)
    pr = REF.mozyme_eimp(ms.numat, ms.norbs, ms.nfirst, ms.nlast, ms.nijbo, ms.mpack, f, p)
    pg = p.copy()
    L_.check(L_.load().mhip_mozyme_eimp(gpu_ctx.h, ms.numat, L_.ptr(ms.nfirst.astype(np.int32)),
                                        L_.ptr(ms.nlast.astype(np.int32)),
                                        L_.ptr(np.asfortranarray(ms.nijbo.astype(np.int32))), ms.mpack,
                                        L_.ptr(f), L_.ptr(pg)))
    assert not np.array_equal(pr, p)
    assert np.array_equal(pg, pr)


def test_fp32_elements_track_fp64(gpu_ctx, lewis_helix_lmos):
    """The mixed form's diagg1 (occupied-virtual elements from FP32 products
    and sums of the FP64 data): on a fresh state the same list as FP64, each
    element and each virtual LMO energy within FP32 rounding of the FP64
    one."""
    L = lewis_helix_lmos
    g64 = mozyme.GpuMozyme(gpu_ctx, L)
    g32 = mozyme.GpuMozyme(gpu_ctx, L, fp32=True)
    r64 = g64.diagg1(L.fao, L.p, 1)
    r32 = g32.diagg1(L.fao, L.p, 1)
    assert r32["nij"] == r64["nij"] and np.array_equal(r32["ifmo"], r64["ifmo"])
    scale = np.max(np.abs(r64["fmo"]))
    err = np.max(np.abs(r32["fmo"] - r64["fmo"]))
    print(f"nij {r64['nij']}, max |fmo| {scale:.3g}, FP32 error {err:.3g}", flush=True)
    assert 0.0 < err < 1e-5 * scale
    # the virtual LMO energies come from the same FP32 products
    assert np.max(np.abs(r32["eigv"] - r64["eigv"])) < 1e-5 * np.max(np.abs(r64["eigv"]))


@pytest.mark.parametrize("nres", [12, 30])
def test_mixed_form_converges_to_the_fp64_heat(gpu_ctx, nres):
    """The whole MOZYME SCF with diagg1's elements in FP32 ends within the
    metric's 1e-3 kcal/mol of the FP64 SCF at SCFCRT 1e-3."""
    nat, xyz = geometry.glycine_helix(nres, 3)
    ms = mozyme.MozymeSystem.from_geometry(nat, xyz)
    ms.upload(gpu_ctx)
    f0 = ms.fock2z(gpu_ctx, ms.diagonal_density())
    lmos = mozyme.lewis_lmos(ms, f0)
    r64 = ms.scf(gpu_ctx, lmos, scfcrt=1e-3)
    r32 = ms.scf(gpu_ctx, lmos, scfcrt=1e-3, fp32_fmo=True)
    print(f"Gly{nres}: FP64 {r64['escf']:.6f} ({r64['niter']} it), mixed {r32['escf']:.6f} ({r32['niter']} it)",
          flush=True)
    assert r64["converged"] and r32["converged"]
    assert abs(r32["escf"] - r64["escf"]) < 1e-3
