"""GPU tests of the accelerated forms: the split-BF16 contraction against FP64,
and the fast SCF (block pseudo-diagonalization, fragment start, split-BF16
while far from convergence) against the plain FP64 SCF on the same system.
The bar is the metric's: |dHf| < 1e-3 kcal/mol; max|dP| < 1e-4 because each
form stops, like the reference, once PL < PLTEST = 0.05 sqrt(SCFCRT) = 5e-5
at SCFCRT = 1e-6, at a different point of its own trajectory."""
import numpy as np
import pytest

from mopac_gpu_amd import geometry, hamiltonian
from mopac_gpu_amd import lib as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_split_bf16_gemm_error(ta, tb):
    lib = L.load()
    rng = np.random.default_rng(3)
    m, n, k = 150, 130, 300
    a = np.asfortranarray(rng.standard_normal((k, m) if ta == "T" else (m, k)))
    b = np.asfortranarray(rng.standard_normal((n, k) if tb == "T" else (k, n)))
    c0 = np.asfortranarray(rng.standard_normal((m, n)))
    opa = a.T if ta == "T" else a
    opb = b.T if tb == "T" else b
    ref = 0.7 * opa @ opb - 0.3 * c0
    c = c0.copy(order="F")
    lib.mhip_gemm_split_bf16(ta.encode(), tb.encode(), m, n, k, 0.7, L.ptr(a), a.shape[0], L.ptr(b),
                             b.shape[0], -0.3, L.ptr(c), m)
    scale = np.abs(opa) @ np.abs(opb)
    err = np.max(np.abs(c - ref) / (0.7 * scale))
    # three BF16 pieces carry 24 bits; FP32 sums over 32-term chunks
    assert err < 1e-6, err
    assert err > 0.0  # it is not the FP64 path
    # and far better than one BF16 rounding (2^-9)
    assert np.max(np.abs(c - ref)) < 1e-4


@pytest.fixture(scope="module")
def water32(gpu_ctx):
    nat, xyz = geometry.water_cluster(32, 1, shape=(4, 4, 2))
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    plain = gpu_ctx.iter(scfcrt=1e-6)
    return s, geometry.fragments(nat, xyz), plain


def _check(r, plain):
    assert r["converged"]
    assert abs(r["escf"] - plain["escf"]) < 1e-3
    assert np.max(np.abs(r["p"] - plain["p"])) < 1e-4
    assert np.max(np.abs(r["eig"] - plain["eig"])) < 1e-4


def test_block_pseudo_diag_converges_to_plain(gpu_ctx, water32):
    s, frag, plain = water32
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(scfcrt=1e-6, pseudo_diag=2)
    _check(r, plain)
    assert r["n_pseudo_diag"] > 0 and r["n_full_diag"] < plain["n_full_diag"]


def test_fragment_start_converges_to_plain(gpu_ctx, water32):
    s, frag, plain = water32
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(scfcrt=1e-6, pseudo_diag=2, block_start=2, fragment=frag, adaptive_eigen=True)
    _check(r, plain)
    assert r["n_fragment_diag"] == 1
    assert r["n_full_diag"] == 1  # only the exit diagonalization


def test_precision_form_converges_to_plain(gpu_ctx, water32):
    s, frag, plain = water32
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(scfcrt=1e-6, pseudo_diag=2, block_start=2, fragment=frag, adaptive_eigen=True,
                     precision=1)
    _check(r, plain)
    assert r["n_split_iter"] > 0


def test_fragment_map_too_large_falls_back(gpu_ctx):
    nat, xyz = geometry.alkane(6, 2)  # one fragment of 38 orbitals
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    plain = gpu_ctx.iter(scfcrt=1e-6)
    r = gpu_ctx.iter(scfcrt=1e-6, pseudo_diag=2, block_start=2, fragment=geometry.fragments(nat, xyz))
    assert r["n_fragment_diag"] == 0
    _check(r, plain)
