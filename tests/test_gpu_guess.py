"""GPU tests of the device-built learned starting density (mhip_guess_water):
the pair descriptors against the host restatement in guess.features, the
network outputs against the standalone forward pass, the assembled density
against guess.assemble on the same outputs, and the SCF started from it
against the SCF started from the host-built density."""
import ctypes

import numpy as np
import pytest

from mopac_gpu_amd import geometry, guess, hamiltonian
from mopac_gpu_amd import lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model():
    m = guess.load_default()
    if m is None:
        pytest.skip("no trained model at mopac_gpu_amd/models/water_guess.npz")
    yield m
    if m._mlp is not None:
        m._mlp.close()


def _host_reference(model, s, info):
    """Host descriptors and the host assembly of the device's own network outputs."""
    a, b, mol, waters = info["a"], info["b"], info["mol"], info["waters"]
    _, _, tree = guess.pair_list(s.nat, s.xyz, mol)
    X, typ = guess.features(s.nat, s.xyz, mol, waters, a, b, tree)
    xn = ((X - model.x_mean) / model.x_std).astype(np.float32)
    y = info["y"] * model.y_std[typ] + model.y_mean[typ]
    return xn, guess.assemble(s, y, a, b, mol, waters)


@pytest.mark.parametrize("nw,seed", [(24, 3), (512, 1)])
def test_device_guess_matches_host(gpu_ctx, model, nw, seed):
    nat, xyz = geometry.water_cluster(nw, seed)
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    print(f"device guess (H2O){nw}", flush=True)
    info = model.to_device(gpu_ctx, s, want_xy=True)
    assert info is not None
    xn, p_host = _host_reference(model, s, info)
    x = info["x"]
    assert x.shape == xn.shape
    assert np.max(np.abs(x - xn) / np.maximum(1.0, np.abs(xn))) < 1e-5
    y_again = model.mlp().forward(x)
    assert np.array_equal(y_again, info["y"])
    p = np.zeros(s.mpack)
    L.check(gpu_ctx.lib.mhip_guess_density(gpu_ctx.h, L.ptr(p)))
    assert np.max(np.abs(p - p_host)) < 1e-12
    # 8 valence electrons per molecule
    d = np.arange(s.norbs)
    assert abs(np.sum(p[d * (d + 1) // 2 + d]) - 8.0 * nw) < 1e-9 * nw


def test_scf_from_device_guess_matches_host_start(gpu_ctx, model):
    nat, xyz = geometry.water_cluster(24, 3)
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    model.to_device(gpu_ctx, s)
    p0 = np.zeros(s.mpack)
    L.check(gpu_ctx.lib.mhip_guess_density(gpu_ctx.h, L.ptr(p0)))
    r_dev = gpu_ctx.iter(scfcrt=1e-6, start_from_guess=True)
    r_host = gpu_ctx.iter(scfcrt=1e-6, p_init=p0)
    plain = gpu_ctx.iter(scfcrt=1e-6)
    assert r_dev["converged"] and r_host["converged"]
    assert r_dev["niter"] == r_host["niter"]
    assert abs(r_dev["escf"] - r_host["escf"]) < 1e-9
    assert abs(r_dev["escf"] - plain["escf"]) < 1e-3
    assert r_dev["niter"] < plain["niter"]


def test_device_guess_needs_upload_and_water(gpu_ctx, model):
    nat, xyz = geometry.alkane(3, 1)
    s = hamiltonian.build(nat, xyz)
    gpu_ctx.upload(s)
    assert model.to_device(gpu_ctx, s) is None  # not water: standard guess
    with pytest.raises(RuntimeError, match="start_from_guess"):
        gpu_ctx.iter(scfcrt=1e-6, start_from_guess=True)
