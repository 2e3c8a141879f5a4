"""mhip_helect_noalloc (the allocation-free variant of mhip_helect) through
the C ABI: device inputs the test allocates itself, the variant's fixed
per-GPU scratch, output checked against mhip_helect on the same inputs
(same kernels, so bit-equal) and against the CPU oracle, on two calls in a
row with the second smaller, plus the size limit."""
import ctypes

import numpy as np
import pytest

from mopac_gpu_amd import geometry, hamiltonian
from mopac_gpu_amd import lib as L
from mopac_gpu_amd.scf import Context
from oracle import oracle

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyKind


def _hip():
    """The HIP runtime libmopachip itself links (already loaded by lib.load())."""
    L.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class _Dev:
    """A device copy of a float64 array, freed on exit."""

    def __init__(self, hip, a):
        self.hip = hip
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), max(8, self.a.nbytes)) == 0
        if self.a.size:
            assert hip.hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def read(self, n):
        out = np.empty(n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _err():
    return (L.load().mhip_last_error() or b"").decode()


def _inputs(k, seed):
    nat, xyz = geometry.water_cluster(k, seed)
    s = hamiltonian.build(nat, xyz)
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((s.norbs, s.norbs)))
    p = oracle.density(np.asfortranarray(q), s.nclose, 2.0)
    pa = 0.5 * p
    f = oracle.fock2(s, s.h, p, pa)
    return s, pa, f


def test_helect_noalloc_matches_helect():
    lib = L.load()
    hip = _hip()
    # (H2O)32 (n = 192), then (H2O)8 (n = 48): the second call reuses the
    # scratch that still holds the first call's 192 partial sums.
    for k, seed in ((32, 7), (8, 9)):
        s, pa, f = _inputs(k, seed)
        with Context(0) as ctx:
            ctx.upload(s)
            e_ref = ctx.helect(pa, s.h, f)
        e_oracle = oracle.helect(s.norbs, pa, s.h, f)
        bufs = [_Dev(hip, x) for x in (pa, s.h, f, np.zeros(1))]
        try:
            print(f"mhip_helect_noalloc: (H2O){k}, norbs {s.norbs}", flush=True)
            rc = lib.mhip_helect_noalloc(s.norbs, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, None)
            assert rc == 0, _err()
            assert hip.hipDeviceSynchronize() == 0
            e = bufs[3].read(1)[0]
        finally:
            for b in bufs:
                b.free()
        assert e == e_ref, (e, e_ref)
        assert abs(e - e_oracle) <= 1e-12 * max(1.0, abs(e_oracle)), (e, e_oracle)


def test_helect_noalloc_rejects_oversize():
    lib = L.load()
    hip = _hip()
    nmax = lib.mhip_helect_noalloc_max_rows()
    assert nmax == 65536
    bufs = [_Dev(hip, np.zeros(1)) for _ in range(4)]
    try:
        # rejected by the size check before any launch
        rc = lib.mhip_helect_noalloc(nmax + 1, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, None)
    finally:
        for b in bufs:
            b.free()
    assert rc != 0
    assert "exceeds the maximum" in _err()
