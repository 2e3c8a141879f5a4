"""GPU tests of the batched SCF (one captured graph for many molecules):
every molecule's heat of formation, iteration count and density against the
CPU oracle running the reference's loop on the same molecule."""
import numpy as np
import pytest

from mopac_gpu_amd import batch as B
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    mols = B.workload(48, seed=11)
    systems = B.build_systems(mols)
    b = B.Batch(systems, scfcrt=1e-6, tiles=[64, 128, 256, 256])
    yield systems, b
    b.close()


def test_batch_matches_oracle_per_molecule(small_batch):
    systems, b = small_batch
    sec, nconv = b.run(iters_per_graph=20, max_launches=20)
    print(f"batch of {len(systems)}: {sec:.3f} s, {nconv} converged", flush=True)
    escf, niter, conv = b.results()
    assert conv.all()
    for m in range(0, len(systems), 3):
        s = systems[m]
        o = oracle.iterate(s, scfcrt=1e-6)
        assert o["converged"]
        assert abs(escf[m] - o["escf"]) < 1e-4, (m, s.norbs, escf[m], o["escf"])
        assert int(niter[m]) == o["niter"], (m, niter[m], o["niter"])
        assert np.max(np.abs(b.density(m) - o["p"])) < 1e-6


def test_batch_is_replayable_and_tiles_do_not_change_results(small_batch):
    systems, b = small_batch
    b.set_tiles([64, 128, 256, 256])
    b.run(iters_per_graph=20, max_launches=20)
    e1, n1, _ = b.results()
    b.set_tiles([128, 64, 128, 512])
    b.run(iters_per_graph=20, max_launches=20)
    e2, n2, _ = b.results()
    assert np.max(np.abs(e1 - e2)) < 1e-6
    assert np.array_equal(n1, n2)


def test_lds_and_global_eigensolver_forms_agree(small_batch):
    """The packed-LDS eigensolver (A in LDS; C too for the classes where both
    fit) against the global-memory form of the same rotations: the same
    iterations, heats within 1e-6 kcal/mol, densities within 1e-7."""
    systems, b = small_batch
    norbs = np.array([s.norbs for s in systems])
    assert norbs.max() > 128 and norbs.min() <= 64  # both LDS forms run
    b.set_tiles([64, 128, 256, 256])
    b.set_eig_lds(False)
    b.run(iters_per_graph=20, max_launches=20)
    eg, ng, cg = b.results()
    pg = [b.density(m) for m in range(len(systems))]
    b.set_eig_lds(True)
    b.run(iters_per_graph=20, max_launches=20)
    el, nl, cl = b.results()
    assert cg.all() and cl.all()
    assert np.array_equal(ng, nl)
    assert np.max(np.abs(eg - el)) < 1e-6
    for m in range(len(systems)):
        assert np.max(np.abs(b.density(m) - pg[m])) < 1e-7, (m, norbs[m])


def test_eigensolver_profile_runs_the_same_iterations(small_batch):
    """The eager, event-timed pass (mhip_batch_eig_profile) runs the same
    iterations as the graph: the same heats and iteration counts, time and
    flops in every populated size class, and no flops where no rotation ran."""
    systems, b = small_batch
    b.set_tiles([64, 128, 256, 256])
    b.run(iters_per_graph=20, max_launches=20)
    e1, n1, _ = b.results()
    ms, fl, rot = b.eig_profile(20)
    e2, n2, c2 = b.results()
    assert c2.all() and np.array_equal(n1, n2)
    assert np.max(np.abs(e1 - e2)) < 1e-9
    norbs = np.array([s.norbs for s in systems])
    for c, (lo, hi) in enumerate([(0, 32), (33, 64), (65, 128), (129, 256)]):
        members = (norbs >= lo) & (norbs <= hi)
        if members.any():
            assert ms[c] > 0 and fl[c] > 0 and rot[c] > 0
            # at least the density and 12 n per rotation of every molecule
            assert fl[c] >= 12.0 * rot[c] * norbs[members].min()
        else:
            assert fl[c] == 0 and rot[c] == 0
