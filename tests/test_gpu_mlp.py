"""GPU tests of the learned starting density: the device MLP (forward,
backward, Adam) against its FP64 CPU reference, and the trained model's
effect on the SCF (fewer iterations, same converged answer)."""
import numpy as np
import pytest

from mopac_gpu_amd import geometry, guess, hamiltonian
from oracle import mlp_ref

pytestmark = pytest.mark.gpu

DIMS = guess.layer_dims()  # [NIN, 256, 256, 256, 16]


def _rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(len(DIMS) - 1):
        parts.append(rng.standard_normal(DIMS[i] * DIMS[i + 1]) * scale / np.sqrt(DIMS[i]))
        parts.append(rng.standard_normal(DIMS[i + 1]) * 0.1)
    return np.concatenate(parts).astype(np.float32)


@pytest.fixture(scope="module")
def mlp():
    m = guess.MLP(seed=0)
    yield m
    m.close()


@pytest.mark.parametrize("batch", [1, 37, 300])
def test_forward_matches_reference(mlp, batch):
    rng = np.random.default_rng(batch)
    params = _rand_params(7)
    mlp.set_params(params)
    x = rng.standard_normal((batch, DIMS[0])).astype(np.float32)
    y = mlp.forward(x)
    y_ref = mlp_ref.forward(params, DIMS, x)[-1]
    # FP32 accumulation against FP64 with the same BF16 rounding of weights and
    # layer inputs; an activation that lands on a BF16 rounding boundary may
    # round the other way (one BF16 ulp = 2^-8 relative).
    scale = np.max(np.abs(y_ref))
    assert np.max(np.abs(y - y_ref)) < 2e-2 * scale
    assert np.sqrt(np.mean((y - y_ref) ** 2)) < 2e-3 * scale


def test_gradients_match_reference(mlp):
    rng = np.random.default_rng(5)
    params = _rand_params(8)
    mlp.set_params(params)
    x = rng.standard_normal((200, DIMS[0])).astype(np.float32)
    t = rng.standard_normal((200, DIMS[-1])).astype(np.float32)
    g, loss = mlp.gradients(x, t)
    loss_ref, g_ref = mlp_ref.loss_and_grad(params, DIMS, x, t)
    assert abs(loss - loss_ref) < 1e-3 * loss_ref
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel < 1e-2
    # each layer's weight and bias block separately (a transposed or shifted
    # block would pass a global norm test only by accident)
    off = 0
    for i in range(len(DIMS) - 1):
        for size in (DIMS[i] * DIMS[i + 1], DIMS[i + 1]):
            blk, blk_ref = g[off:off + size], g_ref[off:off + size]
            assert np.linalg.norm(blk - blk_ref) < 2e-2 * np.linalg.norm(blk_ref) + 1e-9
            off += size


def test_adam_step_matches_numpy(mlp):
    rng = np.random.default_rng(9)
    params = _rand_params(10)
    x = rng.standard_normal((128, DIMS[0])).astype(np.float32)
    t = rng.standard_normal((128, DIMS[-1])).astype(np.float32)
    mlp.set_params(params)
    g, _ = mlp.gradients(x, t)
    lr = 1e-3
    mlp.train_step(x, t, lr)
    after = mlp.get_params()
    # first Adam step: m/(1-b1) = g, v/(1-b2) = g^2  =>  w -= lr * g / (|g| + eps)
    expect = params - lr * g / (np.abs(g) + 1e-8)
    assert np.max(np.abs(after - expect)) < 1e-6 + 1e-3 * lr


def test_training_reduces_loss(mlp):
    rng = np.random.default_rng(11)
    teacher = _rand_params(12, scale=0.5)
    x = rng.standard_normal((512, DIMS[0])).astype(np.float32)
    t = mlp_ref.forward(teacher, DIMS, x)[-1].astype(np.float32)
    mlp.set_params(_rand_params(13))
    first = mlp.train_step(x, t, 1e-3)
    for _ in range(150):
        last = mlp.train_step(x, t, 1e-3)
    assert last < 0.2 * first


def test_learned_guess_saves_iterations(gpu_ctx):
    model = guess.load_default()
    if model is None:
        pytest.skip("no trained model at mopac_gpu_amd/models/water_guess.npz (python -m mopac_gpu_amd.train_guess)")
    nat, xyz = geometry.water_cluster(32, 77)
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    r0 = gpu_ctx.iter(scfcrt=1e-6)
    p0 = model.predict_density(s)
    assert p0 is not None
    # the guess carries the right electron count
    d = np.array([i * (i + 1) // 2 + i for i in range(s.norbs)])
    assert abs(p0[d].sum() - 2 * s.nclose) < 1e-6
    r1 = gpu_ctx.iter(scfcrt=1e-6, p_init=p0)
    assert r1["converged"] and r0["converged"]
    assert r1["niter"] < r0["niter"]
    assert abs(r1["escf"] - r0["escf"]) < 1e-3
    assert np.max(np.abs(r1["p"] - r0["p"])) < 1e-4


def test_guess_falls_back_for_non_water():
    model = guess.load_default()
    if model is None:
        pytest.skip("no trained model at mopac_gpu_amd/models/water_guess.npz")
    nat, xyz = geometry.alkane(4, 1)
    s = hamiltonian.build(nat, xyz)
    assert model.predict_density(s) is None


def test_gradients_match_reference_hcno_dims():
    """The same kernels at the H, C, N, O model's layer sizes (168 inputs)."""
    from mopac_gpu_amd import guess_hcno as H
    dims = H.layer_dims()
    rng = np.random.default_rng(15)
    params = guess.init_params(3, dims)
    m = guess.MLP(params, dims=dims)
    try:
        x = rng.standard_normal((150, dims[0])).astype(np.float32)
        t = rng.standard_normal((150, dims[-1])).astype(np.float32)
        g, loss = m.gradients(x, t)
        loss_ref, g_ref = mlp_ref.loss_and_grad(params, dims, x, t)
        assert abs(loss - loss_ref) < 1e-3 * loss_ref
        assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-2
        y = m.forward(x)
        y_ref = mlp_ref.forward(params, dims, x)[-1]
        assert np.sqrt(np.mean((y - y_ref) ** 2)) < 2e-3 * np.max(np.abs(y_ref))
    finally:
        m.close()


@pytest.mark.parametrize("name,make", [("Gly12 helix", lambda: geometry.glycine_helix(12, 902)),
                                       ("(H2O)20", lambda: geometry.water_cluster(20, 904))])
def test_hcno_guess_saves_iterations(gpu_ctx, name, make):
    """The H, C, N, O model on held-out systems (seeds outside its training
    stream): the plain FP64 SCF from the learned density takes fewer
    iterations than from the standard guess and reaches the same answer
    (heat within 1e-3 kcal/mol at SCFCRT 1e-6; densities within 1e-6 once
    both are converged to 1e-12: at 1e-10 each run stops once its density
    changes by less than PLTEST = 5e-7, so the two stop up to ~1.5e-6 apart)."""
    from mopac_gpu_amd import guess_hcno as H
    model = H.load_default()
    if model is None:
        pytest.skip("no trained model at mopac_gpu_amd/models/hcno_guess.npz (python -m mopac_gpu_amd.train_guess hcno)")
    nat, xyz = make()
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    p0 = model.predict_density(s)
    d = np.arange(s.norbs) * (np.arange(s.norbs) + 1) // 2 + np.arange(s.norbs)
    assert abs(p0[d].sum() - 2 * s.nclose) < 1e-6
    r0 = gpu_ctx.iter(scfcrt=1e-6)
    r1 = gpu_ctx.iter(scfcrt=1e-6, p_init=p0)
    print(f"{name}: standard guess {r0['niter']} iterations, learned {r1['niter']}", flush=True)
    assert r0["converged"] and r1["converged"]
    assert r1["niter"] < r0["niter"]
    assert abs(r1["escf"] - r0["escf"]) < 1e-3
    t0 = gpu_ctx.iter(scfcrt=1e-12)
    t1 = gpu_ctx.iter(scfcrt=1e-12, p_init=p0)
    assert t0["converged"] and t1["converged"]
    assert np.max(np.abs(t1["p"] - t0["p"])) < 1e-6


def test_hcno_guess_falls_back_for_other_elements():
    from mopac_gpu_amd import guess_hcno as H
    model = H.load_default()
    if model is None:
        pytest.skip("no trained model at mopac_gpu_amd/models/hcno_guess.npz")
    assert not H.supported(np.array([8, 1, 1, 16]))

    class S:  # a system with a sulfur atom
        nat = np.array([16, 1, 1])
        xyz = np.zeros((3, 3))
    assert model.predict_density(S()) is None
