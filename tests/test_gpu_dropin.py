"""The reference-boundary drop-ins not covered elsewhere, through the C ABI:
device selection (getGPUInfo / setDevice signatures), the ordered rotation
kernel behind call_rot_cuda_gpu against the oracle's diag, and the
idempotent teardown."""
import ctypes

import numpy as np
import pytest

from mopac_gpu_amd import geometry, hamiltonian
from mopac_gpu_amd import lib as L
from oracle import oracle

pytestmark = pytest.mark.gpu


def _unpack(fp, n):
    f = np.zeros((n, n))
    il = np.tril_indices(n)
    # packed lower triangle by rows: (i, j <= i) at i(i+1)/2 + j
    order = np.argsort(il[0] * (il[0] + 1) // 2 + il[1])
    f[il[0][order], il[1][order]] = fp
    return f + np.tril(f, -1).T


def test_get_gpu_info_and_set_device():
    lib = L.load()
    has = ctypes.c_bool()
    has_double = (ctypes.c_bool * 6)()
    nd = ctypes.c_int()
    names = (ctypes.c_char * 256 * 6)()
    name_size = (ctypes.c_int * 6)()
    mem = (ctypes.c_size_t * 6)()
    clock = (ctypes.c_int * 6)()
    major = (ctypes.c_int * 6)()
    minor = (ctypes.c_int * 6)()
    lib.mhip_getGPUInfo(ctypes.byref(has), has_double, ctypes.byref(nd), names, name_size, mem, clock, major,
                        minor)
    assert has.value and nd.value >= 1
    name = bytes(names[0]).split(b"\0")[0].decode()
    assert len(name) == name_size[0] > 0
    assert "gfx950" in name or "MI355" in name, name
    assert mem[0] > 200 * 2**30          # MI355X: 288 GB of HBM3E
    assert major[0] == 9 and has_double[0]
    for i in range(nd.value, 6):
        assert name_size[i] == 0 and mem[i] == 0 and not has_double[i]
    stat = ctypes.c_bool()
    lib.mhip_setDevice(0, ctypes.byref(stat))
    assert stat.value and lib.mhip_selected_device() == 0
    lib.mhip_setDevice(nd.value + 5, ctypes.byref(stat))
    assert not stat.value and lib.mhip_selected_device() == 0
    lib.mhip_set_mozyme_gpu_pair(0, 0)


def test_call_rot_matches_oracle_diag():
    """call_rot_cuda_gpu's contract: FMO = C_virt^T F C_occ (nvirt x nocc,
    column-major, diag_for_GPU.F90:116-131), tiny = 0.05 max|FMO|, the ordered
    2x2 rotations of diag.F90 applied to C."""
    lib = L.load()
    nat, xyz = geometry.water_cluster(8, 21)
    s = hamiltonian.build(nat, xyz)
    o = oracle.iterate(s, pulay=False, scfcrt=1e-3)
    rng = np.random.default_rng(4)
    p = o["p"] + 1e-3 * rng.standard_normal(o["p"].shape)
    fp = oracle.fock2(s, s.h, p, 0.5 * p)
    n, nocc = s.norbs, s.nclose
    c = np.asfortranarray(o["c"])
    fmo = np.asfortranarray(c[:, nocc:].T @ _unpack(fp, n) @ c[:, :nocc])  # nvirt x nocc
    tiny = 0.05 * np.max(np.abs(fmo))
    v = c.copy(order="F")
    print(f"mhip_call_rot: n {n}, nocc {nocc}", flush=True)
    lib.mhip_call_rot(L.ptr(fmo), L.ptr(np.ascontiguousarray(o["eig"])), L.ptr(v), None, None, nocc, nocc + 1, n,
                      1.5e-7, tiny)
    v_ref, nr = oracle.diag(fp, c, nocc, o["eig"])
    assert nr > 0
    assert np.max(np.abs(v - v_ref)) < 1e-11


def test_destroy_resources_is_idempotent():
    lib = L.load()
    rng = np.random.default_rng(7)
    n = 64
    a = rng.standard_normal((n, n))
    a = np.asfortranarray(a + a.T)
    for _ in range(2):
        aa = a.copy(order="F")
        w = np.zeros(n)
        info = ctypes.c_int()
        lib.mhip_dsyevd(n, L.ptr(aa), n, L.ptr(w), ctypes.byref(info))
        assert info.value == 0
        assert np.max(np.abs(w - np.linalg.eigvalsh(a))) < 1e-10
        lib.mhip_destroy_resources()
        lib.mhip_destroy_resources()  # a second call finds nothing to free


def test_scf_restarts_from_density_file(gpu_ctx, tmp_path):
    """DENOUT then OLDENS: the converged alpha density written as a .den file
    and read back starts the SCF at convergence (p = 2 pa, pold = p)."""
    lib = L.load()
    nat, xyz = geometry.water_cluster(32, 1)
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    # a tight criterion, so the two runs stop at the same density to 1e-6
    r = gpu_ctx.iter(scfcrt=1e-8)
    for formatted in (0, 1):
        path = str(tmp_path / f"w{formatted}.den").encode()
        pa = 0.5 * r["p"]
        assert lib.mhip_den_write(path, s.norbs, s.numat, L.ptr(pa), formatted) == 0
        back = np.zeros(s.mpack)
        assert lib.mhip_den_read(path, s.norbs, s.numat, L.ptr(back), formatted) == 0
        r2 = gpu_ctx.iter(scfcrt=1e-8, p_init=2.0 * back)
        assert r2["converged"] and r2["niter"] < r["niter"]
        assert abs(r2["escf"] - r["escf"]) < 1e-6
        assert np.max(np.abs(r2["p"] - r["p"])) < 1e-6
