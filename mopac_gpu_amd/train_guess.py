"""Train the learned-guess MLP on densities the GPU SCF itself converges.

    python -m mopac_gpu_amd.train_guess [--clusters 240] [--steps 4000] [--out PATH]

1. Seeded water clusters (8-48 molecules, lattice spacing 2.9-3.3 A, jitter
   0.15-0.25 A) are converged with the FP64 GPU SCF (SCFCRT=1e-6).
2. Every atom pair within 4 A (plus self and same-molecule pairs) becomes one
   training sample: features -> its density block in the molecule frames.
3. The MLP is trained on the GPU with the library's own forward/backward
   kernels and Adam; held-out clusters measure the error of the predicted
   density against the converged SCF and the iterations it saves.
Outputs are standardized per pair type (self O, self H, intra O-H, intra H-H,
inter O-O, O-H, H-O, H-H); the learning rate follows a cosine schedule.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import geometry, guess, hamiltonian
from .scf import Context


def cluster(seed):
    rng = np.random.default_rng(10_000 + seed)
    k = int(rng.integers(8, 49))
    spacing = float(rng.uniform(2.9, 3.3))
    jitter = float(rng.uniform(0.15, 0.25))
    return geometry.water_cluster(k, 50_000 + seed, spacing=spacing, jitter=jitter)


def samples_for(ctx, nat, xyz):
    s = hamiltonian.build(nat, xyz)
    ctx.upload(s)
    r = ctx.iter(scfcrt=1e-6)
    mol, waters = guess.find_waters(nat, xyz)
    a, b, tree = guess.pair_list(nat, xyz, mol)
    X, typ = guess.features(nat, xyz, mol, waters, a, b, tree)
    Y = guess.blocks_from_density(s, r["p"], a, b, mol, waters)
    return s, r, X, Y


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--clusters", type=int, default=600)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=guess.DEFAULT_MODEL)
    ap.add_argument("--report", default="")
    args = ap.parse_args(argv)
    t0 = time.time()
    ctx = Context(0)
    Xs, Ys = [], []
    nconv = 0
    for c in range(args.clusters):
        nat, xyz = cluster(c)
        _, r, X, Y = samples_for(ctx, nat, xyz)
        if not r["converged"]:
            continue
        nconv += 1
        Xs.append(X)
        Ys.append(Y)
    X = np.concatenate(Xs).astype(np.float64)
    Y = np.concatenate(Ys).astype(np.float64)
    T = np.argmax(X[:, :8], axis=1)
    t_data = time.time() - t0
    print(f"data: {nconv} converged clusters, {len(X)} pair samples, {t_data:.1f} s", flush=True)
    x_mean = X.mean(axis=0)
    x_std = np.maximum(X.std(axis=0), 1e-3)
    # per pair type: the blocks of different types live on different scales
    y_mean = np.zeros((8, guess.NOUT))
    y_std = np.ones((8, guess.NOUT))
    for t in range(8):
        m = T == t
        if m.any():
            y_mean[t] = Y[m].mean(axis=0)
            y_std[t] = np.maximum(Y[m].std(axis=0), 1e-6)
    Xn = ((X - x_mean) / x_std).astype(np.float32)
    Yn = ((Y - y_mean[T]) / y_std[T]).astype(np.float32)
    mlp = guess.MLP(seed=1)
    rng = np.random.default_rng(2)
    t1 = time.time()
    losses = []
    for step in range(args.steps):
        idx = rng.integers(0, len(Xn), size=args.batch)
        lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * step / max(args.steps, 1)))
        loss = mlp.train_step(Xn[idx], Yn[idx], lr)
        if step % 2000 == 0 or step == args.steps - 1:
            losses.append((step, float(loss)))
            print(f"step {step} loss {loss:.5f}", flush=True)
    t_train = time.time() - t1
    model = guess.GuessModel(mlp.get_params(), x_mean.astype(np.float32), x_std.astype(np.float32),
                             y_mean.astype(np.float32), y_std.astype(np.float32))
    model.save(args.out)
    # held-out evaluation: density error and iterations saved
    evals = []
    for c in range(args.holdout):
        nat, xyz = cluster(100_000 + c)
        s = hamiltonian.build(nat, xyz)
        ctx.upload(s)
        r0 = ctx.iter(scfcrt=1e-6)
        p0 = model.predict_density(s)
        r1 = ctx.iter(scfcrt=1e-6, p_init=p0)
        covered = p0 != 0
        evals.append({"waters": len(nat) // 3, "iter_pdiag": r0["niter"], "iter_learned": r1["niter"],
                      "guess_max_dP_covered": float(np.max(np.abs(p0[covered] - r0["p"][covered]))),
                      "guess_rms_dP_covered": float(np.sqrt(np.mean((p0[covered] - r0["p"][covered]) ** 2))),
                      "dHf": abs(r1["escf"] - r0["escf"]),
                      "max_dP_converged": float(np.max(np.abs(r1["p"] - r0["p"])))})
        print(json.dumps(evals[-1]), flush=True)
    ctx.close()
    rep = {"clusters": nconv, "samples": int(len(X)), "data_s": t_data, "train_s": t_train,
           "steps": args.steps, "batch": args.batch, "losses": losses, "holdout": evals}
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    return rep


if __name__ == "__main__":
    main()
