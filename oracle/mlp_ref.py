"""FP64 CPU reference of the learned-guess MLP — TEST INFRASTRUCTURE.

Same function as the device kernels (k_mlp.hip): BF16-rounded weights and
layer inputs, tanh hidden layers, linear output, MSE loss; gradients with the
BF16 rounding treated as the identity.  Used only to check the GPU forward and
backward passes.
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    special = (a & 0x7F800000) == 0x7F800000
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    r = np.where(special, (a >> 16) << 16, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, dims):
    out, off = [], 0
    for i in range(len(dims) - 1):
        w = params[off:off + dims[i] * dims[i + 1]].reshape(dims[i], dims[i + 1])
        off += dims[i] * dims[i + 1]
        b = params[off:off + dims[i + 1]]
        off += dims[i + 1]
        out.append((w, b))
    return out


def forward(params, dims, x, rnd=None, pdtype=np.float32):
    """Activations of every layer.  rnd (default bf16) rounds weights and layer
    inputs; pdtype is the storage type of the parameters (float32 masters)."""
    rnd = bf16 if rnd is None else rnd
    layers = unpack(np.asarray(params, dtype=pdtype), dims)
    acts = [np.asarray(x, dtype=np.float64)]
    for li, (w, b) in enumerate(layers):
        z = rnd(acts[-1]) @ rnd(w) + b.astype(np.float64)
        acts.append(np.tanh(z) if li + 1 < len(layers) else z)
    return acts


def loss_and_grad(params, dims, x, t, rnd=None, pdtype=np.float32):
    rnd = bf16 if rnd is None else rnd
    layers = unpack(np.asarray(params, dtype=pdtype), dims)
    acts = forward(params, dims, x, rnd, pdtype)
    y = acts[-1]
    t = np.asarray(t, dtype=np.float64)
    ny = y.size
    loss = np.sum((y - t) ** 2) / ny
    d = 2.0 * (y - t) / ny
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        if li + 1 < len(layers):
            d = d * (1.0 - acts[li + 1] ** 2)
        gw = rnd(acts[li]).T @ d
        gb = d.sum(axis=0)
        grads[li] = (gw, gb)
        if li > 0:
            d = d @ rnd(layers[li][0]).T
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat
