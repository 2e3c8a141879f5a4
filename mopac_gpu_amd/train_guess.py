"""Train the learned-guess MLPs on densities the GPU SCF itself converges.

    python -m mopac_gpu_amd.train_guess [--clusters 240] [--steps 4000] [--out PATH]
    python -m mopac_gpu_amd.train_guess hcno [--batches 5] [--steps 40000] [--out PATH]

The first form trains the water model (guess.py), the second the H, C, N, O
model (guess_hcno.py) on helix fragments and water clusters converged by the
batched SCF and longer helices converged one by one; both evaluate held-out
systems at the end.

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




# ---- the H, C, N, O model (guess_hcno) ---------------------------------------

def hcno_workload(nmol, seed):
    """Seeded training molecules: poly-glycine helix fragments of 1-13
    residues (25-253 basis functions) and water clusters of 2-30 molecules,
    alternating."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(nmol):
        if m % 2 == 0:
            out.append(geometry.glycine_helix(int(rng.integers(1, 14)), int(rng.integers(0, 2 ** 31))))
        else:
            out.append(geometry.water_cluster(int(rng.integers(2, 31)), int(rng.integers(0, 2 ** 31))))
    return out


def hcno_charge_table(systems, densities):
    """Mean Mulliken charge per bonding type over converged densities."""
    from . import guess_hcno as H
    acc = {}
    for s, p in zip(systems, densities):
        nbrs, _ = H.bonded_neighbours(s.nat, s.xyz)
        for t, q in zip(H.atom_types(s.nat, nbrs), H.charges_from_density(s, p)):
            a = acc.setdefault(int(t), [0.0, 0])
            a[0] += q
            a[1] += 1
    codes = np.array(sorted(acc), dtype=np.int64)
    return codes, np.array([acc[c][0] / acc[c][1] for c in codes])


def hcno_data(batches, per_batch, pairs_per_mol, seed, log=print):
    """Converged densities from the batched GPU SCF (SCFCRT 1e-6), turned
    into (descriptor, block) samples: up to pairs_per_mol random pairs of
    every converged molecule.  The charge table of the long-range
    descriptors comes from the first batch.  Returns the samples and the
    table."""
    from . import batch as B
    from . import guess_hcno as H
    rng = np.random.default_rng(seed + 7)
    Xs, Ys, Ts = [], [], []
    nconv = 0
    qtab = None
    for k in range(batches):
        t0 = time.time()
        mols = hcno_workload(per_batch, seed + 1000 * k)
        systems = B.build_systems(mols)
        b = B.Batch(systems, scfcrt=1e-6)
        try:
            b.run(iters_per_graph=20, max_launches=50)
            _, _, conv = b.results()
            if qtab is None:
                ok = [m for m in range(len(systems)) if conv[m]]
                qtab = hcno_charge_table([systems[m] for m in ok], [b.density(m) for m in ok])
                log(f"[train] charge table: {len(qtab[0])} bonding types")
            for m, s in enumerate(systems):
                if not conv[m]:
                    continue
                nconv += 1
                prep = H.prepare(s.nat, s.xyz, qtab)
                sel = np.arange(len(prep["a"]))
                if len(sel) > pairs_per_mol:
                    sel = np.sort(rng.choice(len(sel), pairs_per_mol, replace=False))
                Y = H.blocks_from_density(s, b.density(m), prep)
                Xs.append(prep["X"][sel])
                Ys.append(Y[sel].astype(np.float32))
                Ts.append(prep["typ"][sel])
        finally:
            b.close()
        log(f"[train] batch {k}: {per_batch} molecules, {int(conv.sum())} converged, "
            f"{sum(len(x) for x in Xs)} samples so far, {time.time() - t0:.1f} s")
    return Xs, Ys, Ts, nconv, qtab


def hcno_long_data(n_long, pairs_per_helix, seed, qtab, log=print):
    """Helices of 14-48 residues (beyond the batch's 256 basis functions),
    converged one by one with the conventional FP64 GPU SCF (SCFCRT 1e-6):
    the polarization of a long helix's middle, which short fragments lack."""
    from . import guess_hcno as H
    rng = np.random.default_rng(seed + 13)
    Xs, Ys, Ts = [], [], []
    ctx = Context(0)
    nconv = 0
    t0 = time.time()
    try:
        for k in range(n_long):
            nat, xyz = geometry.glycine_helix(int(rng.integers(14, 49)), int(rng.integers(0, 2 ** 31)))
            s = hamiltonian.build(nat, xyz)
            ctx.upload(s)
            r = ctx.iter(scfcrt=1e-6)
            if not r["converged"]:
                continue
            nconv += 1
            prep = H.prepare(nat, xyz, qtab)
            sel = np.arange(len(prep["a"]))
            if len(sel) > pairs_per_helix:
                sel = np.sort(rng.choice(len(sel), pairs_per_helix, replace=False))
            Y = H.blocks_from_density(s, r["p"], prep)
            Xs.append(prep["X"][sel])
            Ys.append(Y[sel].astype(np.float32))
            Ts.append(prep["typ"][sel])
            if k % 20 == 19:
                log(f"[train] long helices: {k + 1} done, {time.time() - t0:.1f} s")
    finally:
        ctx.close()
    return Xs, Ys, Ts, nconv


def train_hcno(argv=None):
    """python -m mopac_gpu_amd.train_guess hcno [--batches 5] [--steps 40000] [--out PATH]"""
    from . import guess_hcno as H
    ap = argparse.ArgumentParser()
    ap.add_argument("--batches", type=int, default=5)
    ap.add_argument("--per-batch", type=int, default=4096)
    ap.add_argument("--pairs-per-mol", type=int, default=160)
    ap.add_argument("--long", type=int, default=120)
    ap.add_argument("--long-pairs", type=int, default=4000)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=H.DEFAULT_MODEL)
    ap.add_argument("--report", default="")
    args = ap.parse_args(argv)
    t0 = time.time()
    Xs, Ys, Ts, nconv, qtab = hcno_data(args.batches, args.per_batch, args.pairs_per_mol, seed=11)
    Xl, Yl, Tl, nlong = hcno_long_data(args.long, args.long_pairs, seed=11, qtab=qtab)
    X, Y, T = np.concatenate(Xs + Xl), np.concatenate(Ys + Yl), np.concatenate(Ts + Tl)
    del Xs, Ys, Ts, Xl, Yl, Tl
    nconv += nlong
    t_data = time.time() - t0
    print(f"[train] data: {nconv} converged densities, {len(X)} pair samples, {t_data:.1f} s", flush=True)
    X = X.astype(np.float64)
    x_mean = X.mean(axis=0)
    x_std = np.maximum(X.std(axis=0), 1e-3)
    y_mean = np.zeros((H.NTYPES, H.NOUT))
    y_std = np.ones((H.NTYPES, H.NOUT))
    for t in range(H.NTYPES):
        m = T == t
        if m.any():
            y_mean[t] = Y[m].mean(axis=0)
            y_std[t] = np.maximum(Y[m].std(axis=0), 1e-6)
    Xn = ((X - x_mean) / x_std).astype(np.float32)
    Yn = ((Y - y_mean[T]) / y_std[T]).astype(np.float32)
    del X
    mlp = guess.MLP(seed=1, dims=H.layer_dims())
    rng = np.random.default_rng(2)
    t1 = time.time()
    losses = []
    for step in range(args.steps):
        idx = rng.integers(0, len(Xn), size=args.batch)
        lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * step / max(args.steps, 1)))
        loss = mlp.train_step(Xn[idx], Yn[idx], lr)
        if step % 5000 == 0 or step == args.steps - 1:
            losses.append((step, float(loss)))
            print(f"[train] step {step} loss {loss:.5f}", flush=True)
    t_train = time.time() - t1
    model = H.HcnoGuessModel(mlp.get_params(), x_mean.astype(np.float32), x_std.astype(np.float32),
                             y_mean.astype(np.float32), y_std.astype(np.float32), qtab[0], qtab[1])
    model.save(args.out)
    mlp.close()
    evals = hcno_evaluate(model)
    rep = {"densities": nconv, "samples": int(len(Xn)), "data_s": t_data, "train_s": t_train,
           "steps": args.steps, "batch": args.batch, "losses": losses, "holdout": evals}
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(rep, fh, indent=1)
    return rep


def hcno_holdout():
    """Held-out systems (seeds outside the training workload's stream)."""
    return [("Gly8 helix", geometry.glycine_helix(8, 901)), ("Gly12 helix", geometry.glycine_helix(12, 902)),
            ("Gly40 helix", geometry.glycine_helix(40, 903)), ("(H2O)20", geometry.water_cluster(20, 904)),
            ("(H2O)64", geometry.water_cluster(64, 905))]


def hcno_evaluate(model):
    """Conventional FP64 GPU SCF of each held-out system from the standard
    guess and from the learned one: iterations, heat difference, guess error."""
    ctx = Context(0)
    out = []
    try:
        for name, (nat, xyz) in hcno_holdout():
            s = hamiltonian.build(nat, xyz)
            ctx.upload(s)
            r0 = ctx.iter(scfcrt=1e-6)
            t = time.time()
            p0 = model.predict_density(s)
            t_guess = time.time() - t
            r1 = ctx.iter(scfcrt=1e-6, p_init=p0)
            cov = p0 != 0
            out.append({"system": name, "norbs": s.norbs, "iter_standard": r0["niter"], "iter_learned": r1["niter"],
                        "guess_s": round(t_guess, 4), "dHf": abs(r1["escf"] - r0["escf"]),
                        "guess_max_dP_covered": float(np.max(np.abs(p0[cov] - r0["p"][cov]))),
                        "guess_rms_dP_covered": float(np.sqrt(np.mean((p0[cov] - r0["p"][cov]) ** 2))),
                        "max_dP_converged": float(np.max(np.abs(r1["p"] - r0["p"])))})
            print("[eval] " + json.dumps(out[-1]), flush=True)
    finally:
        ctx.close()
    return out


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "hcno":
        train_hcno(sys.argv[2:])
    else:
        main()
