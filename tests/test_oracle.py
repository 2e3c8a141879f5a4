"""CPU tests: pin the oracle (oracle/oracle_scf.c) before it is trusted.

1. Routine by routine against MOPAC's own Fortran (fock2/fock1/jab/kab,
   helect, cnvg, pulay, diag, osinv compiled from the reference sources by
   oracle/build_ref.py), on identical inputs.
2. End to end against the reference's published outputs (eigenvalues and
   charges of tests/keywords/1SCF.out, CdotIdot.out; charges of the API test).
3. Against what the mathematics fixes (density = 2 C_occ C_occ^T, helect =
   1/2 tr P(H+F), the eigen-decomposition, a second formulation of fock2,
   Gauss-Jordan inverse, invariance of the SCF energy under rigid motion).
"""
import json
import os

import numpy as np
import pytest

from mopac_gpu_amd import geometry, hamiltonian
from oracle import oracle
from oracle import ref as REF

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_outputs.json")))
needs_ref = pytest.mark.skipif(not REF.available(),
                               reason="oracle/_ref not built (needs the reference sources: oracle/build_ref.py)")


def _system(name):
    g = GOLD[name]
    return hamiltonian.build(np.array(g["nat"]), np.array(g["xyz"]))


def _systems():
    return {"c2h4": _system("c2h4_1scf"), "water4": hamiltonian.build(*geometry.water_cluster(4, 11)),
            "alkane": hamiltonian.build(*geometry.alkane(3, 2)),
            "glycine": hamiltonian.build(*geometry.glycine_helix(1, 3))}


SYSTEMS = _systems()


def _rand_density(s, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((s.norbs, s.norbs)))
    return oracle.density(np.asfortranarray(q), s.nclose)


def _diag_idx(n):
    return np.array([i * (i + 1) // 2 + i for i in range(n)])


def _unpack(p, n):
    m = np.zeros((n, n))
    m[np.tril_indices(n)] = p
    return m + m.T - np.diag(np.diag(m))


# ---- 1. against the reference's own routines --------------------------------

@needs_ref
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_fock2_matches_reference_fortran(name):
    s = SYSTEMS[name]
    p = _rand_density(s, 3)
    pa = 0.5 * p
    pa[_diag_idx(s.norbs)] *= 0.97
    f_o = oracle.fock2(s, s.h, p, pa)
    f_r = REF.fock2(s, s.h, p, pa)
    assert np.max(np.abs(f_o - f_r)) < 1e-12 * np.max(np.abs(f_r))
    assert oracle.helect(s.norbs, pa, s.h, f_o) == pytest.approx(REF.helect(s.norbs, pa, s.h, f_r), rel=1e-14)


@needs_ref
def test_cnvg_matches_reference_fortran():
    s = SYSTEMS["water4"]
    REF.setup(s)
    d = _diag_idx(s.norbs)
    pold = _rand_density(s, 4)
    for it in range(1, 9):
        pnew = 0.3 * _rand_density(s, 10 + it) + 0.7 * pold
        a1, b1, c1 = pnew.copy(), pold.copy(), pold[d].copy()
        a2, b2, c2 = pnew.copy(), pold.copy(), pold[d].copy()
        pl_o = oracle.cnvg(a1, b1, c1, it)
        pl_r = REF.cnvg(a2, b2, c2, it)
        assert pl_o == pl_r
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and np.array_equal(c1, c2)
        pold = a1


@needs_ref
@pytest.mark.parametrize("name", ["c2h4", "water4"])
def test_pulay_sequence_matches_reference_fortran(name):
    s = SYSTEMS[name]
    ncall = 14  # wraps the 10-deep history
    fs = np.stack([oracle.fock2(s, s.h, _rand_density(s, 30 + k), 0.5 * _rand_density(s, 30 + k))
                   for k in range(ncall)])
    ps = np.stack([0.5 * _rand_density(s, 50 + k) for k in range(ncall)])
    REF.setup(s)
    f_o, pl_o = oracle.pulay_seq(s.norbs, fs, ps)
    f_r, pl_r = REF.pulay_seq(s.norbs, fs, ps)
    assert np.max(np.abs(f_o - f_r)) < 1e-12 * np.max(np.abs(f_r))
    assert np.max(np.abs(pl_o - pl_r)) < 1e-13 * np.max(np.abs(pl_r))


@needs_ref
@pytest.mark.parametrize("name", ["c2h4", "water4", "glycine"])
def test_pseudo_diag_matches_reference_fortran(name):
    s = SYSTEMS[name]
    o = oracle.iterate(s, scfcrt=1e-3)
    rng = np.random.default_rng(6)
    p = o["p"] + 1e-3 * rng.standard_normal(s.mpack)
    f = oracle.fock2(s, s.h, p, 0.5 * p)
    REF.setup(s)
    v_o, nrot = oracle.diag(f, o["c"], s.nclose, o["eig"])
    v_r = REF.diag(f, o["c"], s.nclose, o["eig"])
    assert nrot > 0
    assert np.max(np.abs(v_o - v_r)) < 1e-14


@needs_ref
@pytest.mark.parametrize("n", [2, 5, 11])
def test_osinv_matches_reference_fortran(n):
    a = np.random.default_rng(n).standard_normal((n, n))
    inv_o, d_o = oracle.osinv(a)
    inv_r, d_r = REF.osinv(a)
    assert np.array_equal(inv_o, inv_r) and d_o == d_r


# ---- 2. against the reference's published outputs ---------------------------

def _charges(s, p, nat):
    d = _diag_idx(s.norbs)
    core = {1: 1.0, 6: 4.0, 7: 5.0, 8: 6.0}
    return np.array([core[z] - p[d[s.nfirst[a] - 1:s.nlast[a]]].sum() for a, z in enumerate(nat)])


@pytest.mark.parametrize("name", ["co2_1scf", "c2h4_1scf"])
def test_scf_reproduces_published_eigenvalues_and_charges(name):
    g = GOLD[name]
    s = _system(name)
    o = oracle.iterate(s, scfcrt=1e-4)
    assert o["converged"]
    ev = np.array(g["eigenvalues"])
    # printed with 5 decimals
    assert np.max(np.abs(o["eig"][:len(ev)] - ev)) < 6e-6
    assert np.max(np.abs(_charges(s, o["p"], g["nat"]) - np.array(g["charges"]))) < 6e-6


def test_scf_pulay_reproduces_api_charges():
    g = GOLD["h2o_api"]
    s = _system("h2o_api")
    o = oracle.iterate(s, pulay=True, scfcrt=1e-4)
    assert o["converged"]
    assert np.max(np.abs(_charges(s, o["p"], g["nat"]) - np.array(g["charges"]))) < 6e-6
    # with PM7's post-SCF corrections (dispersion; no hydrogen bond in one
    # water) the heat of formation is the API's known answer
    assert abs(o["escf"] + s.e_disp + s.e_hbond - g["hof"]) < 2e-5


# ---- 3. against what the mathematics fixes ----------------------------------

@pytest.mark.parametrize("name", ["c2h4", "glycine"])
def test_fock2_two_formulations_agree(name):
    s = SYSTEMS[name]
    p = _rand_density(s, 8)
    pa = 0.45 * p
    f1 = oracle.fock2(s, s.h, p, pa)
    f2 = oracle.fock2(s, s.h, p, pa, general=True)
    assert np.max(np.abs(f1 - f2)) < 1e-12 * np.max(np.abs(f1))


def test_density_and_energy_closed_forms():
    s = SYSTEMS["water4"]
    rng = np.random.default_rng(2)
    c, _ = np.linalg.qr(rng.standard_normal((s.norbs, s.norbs)))
    c = np.asfortranarray(c)
    p = oracle.density(c, s.nclose, 2.0)
    full = 2.0 * c[:, :s.nclose] @ c[:, :s.nclose].T
    assert np.max(np.abs(_unpack(p, s.norbs) - full)) < 1e-13
    # idempotency of the projector and the electron count
    assert abs(np.trace(full) - 2 * s.nclose) < 1e-10
    assert np.max(np.abs(full @ full - 2.0 * full)) < 1e-12
    f = oracle.fock2(s, s.h, p, 0.5 * p)
    e = oracle.helect(s.norbs, 0.5 * p, s.h, f)
    e_ref = 0.5 * np.sum(_unpack(0.5 * p, s.norbs) * (_unpack(s.h, s.norbs) + _unpack(f, s.norbs)))
    assert e == pytest.approx(e_ref, rel=1e-13)


def test_eigenvectors_match_numpy_and_shift_in_place():
    s = SYSTEMS["glycine"]
    p = _rand_density(s, 5)
    f = oracle.fock2(s, s.h, p, 0.5 * p)
    f0 = f.copy()
    c, e = oracle.eigenvectors(f)
    d = _diag_idx(s.norbs)
    assert np.allclose(f[d] - f0[d], 1e-10 * np.arange(1, s.norbs + 1), rtol=0, atol=1e-13)
    m = _unpack(f, s.norbs)
    assert np.max(np.abs(e - np.linalg.eigvalsh(m))) < 1e-10
    assert np.max(np.abs(m @ c - c * e)) < 1e-10
    assert np.max(np.abs(c.T @ c - np.eye(s.norbs))) < 1e-12


def test_osinv_inverse_and_determinant():
    a = np.random.default_rng(3).standard_normal((7, 7))
    inv, d = oracle.osinv(a)
    assert np.max(np.abs(inv @ a - np.eye(7))) < 1e-12
    assert d == pytest.approx(np.linalg.det(a), rel=1e-12)


def test_scf_energy_invariant_under_rigid_motion():
    nat, xyz = geometry.glycine_helix(1, 4)
    s0 = hamiltonian.build(nat, xyz)
    th = 0.7
    r = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    r = r @ np.array([[1, 0, 0], [0, np.cos(1.1), -np.sin(1.1)], [0, np.sin(1.1), np.cos(1.1)]])
    s1 = hamiltonian.build(nat, xyz @ r.T + np.array([3.0, -2.0, 5.0]))
    e0 = oracle.iterate(s0, scfcrt=1e-9)
    e1 = oracle.iterate(s1, scfcrt=1e-9)
    assert e0["converged"] and e1["converged"]
    assert abs(e0["escf"] - e1["escf"]) < 1e-6
    assert np.max(np.abs(e0["eig"] - e1["eig"])) < 1e-6


# ---- MOZYME: the reference routines driven through the harness -------------

def _helix_lmos(rcut):
    from mopac_gpu_amd import mozyme
    nat, xyz = geometry.glycine_helix(2, 3)
    s = hamiltonian.build(nat, xyz)
    o = oracle.iterate(s, scfcrt=1e-6)
    n = s.norbs
    return s, o, mozyme.from_scf(s, o["c"], _unpack(o["f"], n), _unpack(o["p"], n), s.nclose, keep=1e-3,
                                 rcut=rcut)


@needs_ref
def test_reference_mozyme_density_is_the_lmo_projector():
    from mopac_gpu_amd import mozyme
    s, o, L = _helix_lmos(6.0)
    p = REF.Mozyme(L).density()
    d = np.zeros((s.norbs, s.norbs))
    for k in range(L.nocc):
        x = mozyme.lmo_dense(L, k, True)
        d += 2.0 * np.outer(x, x)
    assert np.max(np.abs(p - mozyme.sparse_from_dense(d, L.nfirst, L.iorbs, L.nijbo, L.mpack))) < 1e-14


@needs_ref
def test_reference_helecz_equals_dense_helect_when_all_pairs_are_stored():
    from mopac_gpu_amd import mozyme
    s, o, L = _helix_lmos(1e3)
    h = mozyme.sparse_from_dense(_unpack(s.h, s.norbs), L.nfirst, L.iorbs, L.nijbo, L.mpack)
    e = REF.Mozyme(L).helecz(L.p, h, L.fao)
    assert e == pytest.approx(oracle.helect(s.norbs, o["p"], s.h, o["f"]), rel=1e-13)


def test_localized_orbitals_span_the_occupied_space():
    from mopac_gpu_amd import mozyme
    s, o, L = _helix_lmos(6.0)
    c = o["c"][:, :s.nclose]
    q = mozyme.pipek_mezey(c, L.nfirst, L.iorbs)
    assert np.max(np.abs(q.T @ q - np.eye(s.nclose))) < 1e-12
    assert np.max(np.abs(q @ q.T - c @ c.T)) < 1e-12


@pytest.mark.skipif(not REF.available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("r_near,r_dipole,feather", [(6.0, 7.0, False), (4.0, 6.0, False), (30.0, 31.0, False),
                                                     (4.0, 6.0, True), (6.0, 8.0, True)])
def test_mozyme_fock2z_matches_reference_fortran(r_near, r_dipole, feather):
    """The restatement of buildf/fock2z (stored pairs through the C fock2,
    point-charge and dipole terms for the others) against the reference's
    own fz2n, SEMIDIRECT, on a poly-glycine helix with all three pair classes."""
    from mopac_gpu_amd import mozyme
    nat, xyz = geometry.glycine_helix(6, 3)
    s = hamiltonian.build(nat, xyz)
    ms = mozyme.MozymeSystem(s, r_near=r_near, r_dipole=r_dipole, feather=feather)
    if r_near < 10:
        assert (ms.nijbo == -2).any() and (ms.nijbo == -1).any()
    rng = np.random.default_rng(1)
    pd = rng.standard_normal((s.norbs, s.norbs)) * 0.1
    pd = pd + pd.T + np.eye(s.norbs)
    p = mozyme.sparse_from_dense(pd, ms.nfirst, ms.nlast - ms.nfirst + 1, ms.nijbo, ms.mpack)
    f_ref = REF.mozyme_fock2z(ms, p, feather=feather)
    f = oracle.mozyme_fock2z(s, ms.nijbo, ms.mpack, ms.h, ms.am, ms.ad, ms.dd, p, feather=feather)
    assert np.max(np.abs(f - f_ref)) < 1e-11 * np.max(np.abs(f_ref))


@pytest.mark.parametrize("r_near,r_dipole,tol", [(30.0, 31.0, 1e-6), (8.0, 9.0, 1e-4), (6.0, 7.0, 0.15)])
def test_mozyme_energy_at_the_conventional_density(r_near, r_dipole, tol):
    """The MOZYME Hamiltonian (mhip_pm7_hamiltonian_mozyme) and sparse Fock
    build (the restated fz2n with PM7's point-charge switch) evaluated at the
    converged conventional density give the conventional heat of formation:
    to 1e-6 kcal/mol when every pair is stored (only the resonance beyond
    MOZYME's 7 A CUTOFS differs); to 1e-4 kcal/mol when the approximated
    pairs lie beyond the 7 A switch (PM7's NDDO terms are point charges there
    already; what remains is the dropped resonance and core-core exponentials);
    within 0.15 kcal/mol when the dipole shell starts at 6 A."""
    import ctypes
    from mopac_gpu_amd import lib as L
    from mopac_gpu_amd import mozyme
    nat, xyz = geometry.glycine_helix(6, 3)
    s = hamiltonian.build(nat, xyz)
    n = s.norbs
    o = oracle.iterate(s, scfcrt=1e-7)
    pd = np.zeros((n, n))
    pd[np.tril_indices(n)] = o["p"]
    pd = pd + pd.T - np.diag(np.diag(pd))
    ms = mozyme.MozymeSystem(s, r_near=r_near, r_dipole=r_dipole, feather=True)
    iorbs = ms.nlast - ms.nfirst + 1
    p = mozyme.sparse_from_dense(pd, ms.nfirst, iorbs, ms.nijbo, ms.mpack)
    f = oracle.mozyme_fock2z(s, ms.nijbo, ms.mpack, ms.h, ms.am, ms.ad, ms.dd, p, feather=True)
    e = 0.0  # helecz: stored blocks, diagonal elements of diagonal blocks at one half
    for a in range(ms.numat):
        for b in range(a + 1):
            k = ms.nijbo[a, b]
            if k < 0:
                continue
            if a == b:
                q = 0
                for i in range(iorbs[a]):
                    for j in range(i + 1):
                        e += (0.5 if i == j else 1.0) * p[k + q] * (ms.h[k + q] + f[k + q])
                        q += 1
            else:
                m = iorbs[a] * iorbs[b]
                e += np.sum(p[k:k + m] * (ms.h[k:k + m] + f[k:k + m]))
    escf = (e + ms.enuclr) * L.load().mhip_ev_to_kcal() + s.atheat
    assert abs(escf - o["escf"]) < tol
