"""The configured workloads at their full sizes on the GPU, against the CPU
oracle with the metric's bar (|dHf| < 1e-3 kcal/mol, max|dP| < 1e-6):
  - (H2O)512 (n = 3072, 2048 occupied): the plain FP64 SCF, equal iteration
    count with the oracle's, and the accelerated form (learned start when a
    trained model is present, block pseudo-diagonalization, split-BF16
    contractions far from convergence);
  - the 4096-molecule batch (alkanes and water clusters, 6-180 basis
    functions) in one captured graph, with a seeded sample of 64 molecules
    and every n = 180 molecule checked against the oracle."""
import numpy as np
import pytest

from mopac_gpu_amd import batch as B
from mopac_gpu_amd import geometry, hamiltonian
from mopac_gpu_amd import guess as learned
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water512():
    nat, xyz = geometry.water_cluster(512, 2, shape=(8, 8, 8))
    s = hamiltonian.build(nat, xyz)
    print(f"(H2O)512: n {s.norbs}, oracle SCF ...", flush=True)
    o = oracle.iterate(s, scfcrt=1e-6)
    assert o["converged"]
    return nat, xyz, s, o


def test_water512_plain_matches_oracle(gpu_ctx, water512):
    nat, xyz, s, o = water512
    gpu_ctx.upload(s)
    r = gpu_ctx.iter(scfcrt=1e-6)
    assert r["converged"]
    assert abs(r["escf"] - o["escf"]) < 1e-3
    assert np.max(np.abs(r["p"] - o["p"])) < 1e-6
    # at n = 3072 the reference algorithm's own iteration count depends on
    # the BLAS thread count of its dsyevd (18 with 8 MKL threads, 20 with 16,
    # on this workload): the trajectory sits near a decision of the stopping
    # test, so the count is held to that spread rather than to equality
    assert abs(r["niter"] - o["niter"]) <= 2


@pytest.mark.parametrize("precision", [0, 1])
def test_water512_accelerated_form_matches_oracle(gpu_ctx, water512, precision):
    nat, xyz, s, o = water512
    gpu_ctx.upload(s)
    model = learned.load_default()
    have = model is not None and model.to_device(gpu_ctx, s) is not None
    r = gpu_ctx.iter(scfcrt=1e-6, adaptive_eigen=True, precision=precision, start_from_guess=have,
                     pseudo_diag=2, block_start=2, fragment=geometry.fragments(nat, xyz), want=("p", "eig", "c"))
    assert r["converged"]
    assert abs(r["escf"] - o["escf"]) < 1e-3
    assert np.max(np.abs(r["p"] - o["p"])) < 1e-6
    if precision:
        assert r["n_split_iter"] > 0
    # the exit eigenvectors are orthonormal to FP64 precision
    c = r["c"]
    g = c.T @ c
    assert np.max(np.abs(g - np.eye(s.norbs))) < 1e-11


def test_full_4096_batch_matches_oracle_sample():
    mols = B.workload(4096, seed=4)
    systems = B.build_systems(mols)
    b = B.Batch(systems, scfcrt=1e-6)
    try:
        sec, nconv = b.run(iters_per_graph=20, max_launches=40)
        print(f"batch of {len(systems)}: {sec:.3f} s, {nconv} converged", flush=True)
        escf, niter, conv = b.results()
        assert conv.all()
        norbs = np.array([x.norbs for x in systems])
        rng = np.random.default_rng(4)
        sample = set(rng.choice(len(systems), 64, replace=False).tolist()) | set(np.flatnonzero(norbs == 180).tolist())
        for m in sorted(sample):
            o = oracle.iterate(systems[m], scfcrt=1e-6)
            assert o["converged"]
            assert abs(escf[m] - o["escf"]) < 1e-3, (m, norbs[m], escf[m], o["escf"])
            assert np.max(np.abs(b.density(m) - o["p"])) < 1e-6, (m, norbs[m])
            assert int(niter[m]) == o["niter"], (m, norbs[m], niter[m], o["niter"])
    finally:
        b.close()
