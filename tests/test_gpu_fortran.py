"""GPU test of the Fortran host side: tests/fortran/test_mopac_hip.F90 calls
libmopachip through the mopac_hip module (mopac_gpu_amd/fortran/mopac_hip.F90)
with the calling surface of fock2, helect, eigenvectors_LAPACK, densit and
iter; every result is checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from mopac_gpu_amd import build as B
from mopac_gpu_amd import geometry, hamiltonian
from oracle import oracle

pytestmark = pytest.mark.gpu


def _write_system(path, s, ptest):
    with open(path, "wb") as fh:
        np.array([s.numat, s.norbs, s.nclose], dtype=np.int32).tofile(fh)
        np.array([len(s.w)], dtype=np.int64).tofile(fh)
        np.array([s.enuclr, s.atheat], dtype=np.float64).tofile(fh)
        np.asarray(s.nfirst, dtype=np.int32).tofile(fh)
        np.asarray(s.nlast, dtype=np.int32).tofile(fh)
        for a in (s.h, s.w, s.pdiag, ptest):
            np.asarray(a, dtype=np.float64).tofile(fh)


def _read_results(path, s):
    mp, n = s.mpack, s.norbs
    raw = open(path, "rb").read()
    off = 0

    def take(cnt, dt):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=cnt, offset=off)
        off += a.nbytes
        return a

    out = {"f": take(mp, np.float64), "e": take(1, np.float64)[0], "fcopy": take(mp, np.float64),
           "eig": take(n, np.float64), "p": take(mp, np.float64), "escf": take(1, np.float64)[0],
           "niter": int(take(1, np.int32)[0]), "converged": int(take(1, np.int32)[0]),
           "pscf": take(mp, np.float64), "eigscf": take(n, np.float64)}
    assert off == len(raw)
    return out


@pytest.mark.skipif(not os.path.exists(B.FORTRAN_TEST_EXE), reason="Fortran host not built (no amdflang)")
def test_fortran_host_matches_oracle(tmp_path):
    nat, xyz = geometry.water_cluster(8, 5)
    s = hamiltonian.build(nat, xyz)
    rng = np.random.default_rng(1)
    q, _ = np.linalg.qr(rng.standard_normal((s.norbs, s.norbs)))
    ptest = oracle.density(np.asfortranarray(q), s.nclose)
    fin, fout = tmp_path / "system.bin", tmp_path / "results.bin"
    _write_system(fin, s, ptest)
    print(f"running {B.FORTRAN_TEST_EXE} on (H2O)8", flush=True)
    r = subprocess.run([B.FORTRAN_TEST_EXE, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g = _read_results(fout, s)
    f_ref = oracle.fock2(s, s.h, ptest, 0.5 * ptest)
    assert np.max(np.abs(g["f"] - f_ref)) < 1e-11 * np.max(np.abs(f_ref))
    assert g["e"] == pytest.approx(oracle.helect(s.norbs, 0.5 * ptest, s.h, f_ref), rel=1e-12)
    # eigenvectors_LAPACK's in-place +i*1e-10 on the diagonal of xmat, nothing else
    d = np.array([i * (i + 1) // 2 + i for i in range(s.norbs)])
    off = np.ones(s.mpack, bool)
    off[d] = False
    assert np.array_equal(g["fcopy"][off], g["f"][off])
    assert np.max(np.abs(g["fcopy"][d] - g["f"][d] - 1e-10 * np.arange(1, s.norbs + 1))) < 1e-14
    fc = f_ref.copy()
    c_ref, e_ref = oracle.eigenvectors(fc)
    assert np.max(np.abs(g["eig"] - e_ref)) < 1e-9
    p_ref = oracle.density(c_ref, s.nclose, 2.0)
    assert np.max(np.abs(g["p"] - p_ref)) < 1e-9
    o = oracle.iterate(s, scfcrt=1e-6)
    assert g["converged"] == 1 and g["niter"] == o["niter"]
    assert abs(g["escf"] - o["escf"]) < 1e-6
    assert np.max(np.abs(g["pscf"] - o["p"])) < 1e-6
    assert np.max(np.abs(g["eigscf"] - o["eig"])) < 1e-6
