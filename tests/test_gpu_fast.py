"""GPU tests of the accelerated forms: the split-BF16 contraction against FP64,
and the fast SCF (block pseudo-diagonalization, fragment start, split-BF16
while far from convergence) against the plain FP64 SCF on the same system.
The bar is the metric's, |dHf| < 1e-3 kcal/mol and max|dP| < 1e-6: at the
metric's SCFCRT = 1e-6 each form stops, like the reference, once PL < PLTEST
= 5e-5, at a different point of its own trajectory, so the density bar is
checked with both forms converged to SCFCRT = 1e-10 (at 1e-8 the two still
differ by their own convergence error, ~1.5e-6 on (H2O)32)."""
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


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(256, 256, 1024), (256, 512, 2048)])
def test_split_bf16_gemm_k_split(ta, tb, m, n, k):
    """Shapes whose padded K lets the launcher split K in 2 or 4 slices (parts
    summed in FP64 after the first slice adds beta C), alpha != 1 and beta != 0
    on a random C."""
    lib = L.load()
    rng = np.random.default_rng(m + n + k)
    a = np.asfortranarray(rng.standard_normal((k, m) if ta == "T" else (m, k)))
    b = np.asfortranarray(rng.standard_normal((n, k) if tb == "T" else (k, n)))
    c0 = np.asfortranarray(rng.standard_normal((m, n)))
    opa = a.T if ta == "T" else a
    opb = b.T if tb == "T" else b
    ref = -1.3 * opa @ opb + 0.6 * c0
    c = c0.copy(order="F")
    lib.mhip_gemm_split_bf16(ta.encode(), tb.encode(), m, n, k, -1.3, L.ptr(a), a.shape[0], L.ptr(b),
                             b.shape[0], 0.6, L.ptr(c), m)
    scale = np.abs(opa) @ np.abs(opb)
    assert np.max(np.abs(c - ref) / (1.3 * scale + 0.6 * np.abs(c0))) < 1e-6


@pytest.fixture(scope="module")
def water32(gpu_ctx):
    nat, xyz = geometry.water_cluster(32, 1, shape=(4, 4, 2))
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    plain = {crt: gpu_ctx.iter(scfcrt=crt) for crt in (1e-6, 1e-10)}
    return s, geometry.fragments(nat, xyz), plain


def _check(r, plain, crt=1e-6):
    assert r["converged"]
    assert abs(r["escf"] - plain[crt]["escf"]) < 1e-3
    if crt <= 1e-10:
        assert np.max(np.abs(r["p"] - plain[crt]["p"])) < 1e-6
        assert np.max(np.abs(r["eig"] - plain[crt]["eig"])) < 1e-6


@pytest.mark.parametrize("crt", [1e-6, 1e-10])
def test_block_pseudo_diag_converges_to_plain(gpu_ctx, water32, crt):
    s, frag, plain = water32
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(scfcrt=crt, pseudo_diag=2)
    _check(r, plain, crt)
    assert r["n_pseudo_diag"] > 0 and r["n_full_diag"] < plain[crt]["n_full_diag"]


@pytest.mark.parametrize("crt", [1e-6, 1e-10])
def test_fragment_start_converges_to_plain(gpu_ctx, water32, crt):
    s, frag, plain = water32
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(scfcrt=crt, pseudo_diag=2, block_start=2, fragment=frag, adaptive_eigen=True,
                     want=("p", "eig", "c"))
    _check(r, plain, crt)
    # the exit eigenvectors (block steps + Bjorck, then the grouped Jacobi
    # exit) are orthonormal to FP64 precision
    c = r["c"]
    assert np.max(np.abs(c.T @ c - np.eye(s.norbs))) < 1e-12
    assert r["n_fragment_diag"] == 1
    assert r["n_full_diag"] == 1  # only the exit diagonalization


@pytest.mark.parametrize("crt", [1e-6, 1e-10])
def test_precision_form_converges_to_plain(gpu_ctx, water32, crt):
    s, frag, plain = water32
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(scfcrt=crt, pseudo_diag=2, block_start=2, fragment=frag, adaptive_eigen=True,
                     precision=1)
    _check(r, plain, crt)
    assert r["n_split_iter"] > 0


def test_fragment_map_too_large_falls_back(gpu_ctx):
    nat, xyz = geometry.alkane(6, 2)  # one fragment of 38 orbitals
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    plain = {1e-6: gpu_ctx.iter(scfcrt=1e-6)}
    r = gpu_ctx.iter(scfcrt=1e-6, pseudo_diag=2, block_start=2, fragment=geometry.fragments(nat, xyz))
    assert r["n_fragment_diag"] == 0
    _check(r, plain)
