"""Batched SCF of many small molecules (mhip_batch_*), the batch workload
generator, and the project's autotuner for the batched eigensolver.

Workload (seed 4): half seeded random alkanes CnH2n+2 (n = 1..20), half
water clusters (H2O)k (k = 1..30, 3.1 A lattice recipe), 8-180 basis
functions each.  Every molecule runs iter's RHF loop (no PULAY) on the
device; the whole batch is one captured hipGraph replayed until every
molecule is done.

Autotuner: the eigensolver (one workgroup per molecule) has a workgroup size
per size class (n <= 32, 64, 128, 256).  `autotune` times the whole batched
SCF for each candidate size of one class at a time (the others fixed), keeps
the fastest, and writes the table to mopac_gpu_amd/tuning/batch_tiles.json,
which `Batch` loads by default.
"""
import ctypes
import json
import os
import time

import numpy as np

from . import geometry, hamiltonian
from . import lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
TILES_JSON = os.path.join(HERE, "tuning", "batch_tiles.json")
CANDIDATES = (64, 128, 256, 512, 1024)


class MoleculeC(ctypes.Structure):
    _fields_ = [("numat", ctypes.c_int), ("norbs", ctypes.c_int), ("nclose", ctypes.c_int),
                ("nfirst", ctypes.c_void_p), ("nlast", ctypes.c_void_p), ("n2elec", ctypes.c_longlong),
                ("h", ctypes.c_void_p), ("w", ctypes.c_void_p), ("pdiag", ctypes.c_void_p),
                ("enuclr", ctypes.c_double), ("atheat", ctypes.c_double), ("scfcrt", ctypes.c_double),
                ("itrmax", ctypes.c_int)]


def workload(nmol=4096, seed=4):
    """(nat, xyz) of the batch: alternating alkanes and water clusters."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for m in range(nmol):
        if m % 2 == 0:
            n = int(rng.integers(1, 21))
            out.append(geometry.alkane(n, int(rng.integers(0, 2 ** 31))))
        else:
            k = int(rng.integers(1, 31))
            out.append(geometry.water_cluster(k, int(rng.integers(0, 2 ** 31))))
    return out


def load_tiles():
    if os.path.exists(TILES_JSON):
        with open(TILES_JSON) as fh:
            return [int(x) for x in json.load(fh)["tiles"]]
    return [64, 128, 256, 256]


class Batch:
    def __init__(self, systems, scfcrt=1e-6, itrmax=2000, tiles=None):
        self.lib = L.load()
        self.systems = systems
        self._keep = []
        mols = (MoleculeC * len(systems))()
        for i, s in enumerate(systems):
            nf = np.ascontiguousarray(s.nfirst, dtype=np.int32)
            nl = np.ascontiguousarray(s.nlast, dtype=np.int32)
            h = np.ascontiguousarray(s.h, dtype=np.float64)
            w = np.ascontiguousarray(s.w, dtype=np.float64)
            pd = np.ascontiguousarray(s.pdiag, dtype=np.float64)
            self._keep += [nf, nl, h, w, pd]
            mols[i] = MoleculeC(numat=s.numat, norbs=s.norbs, nclose=s.nclose, nfirst=nf.ctypes.data,
                                nlast=nl.ctypes.data, n2elec=len(w), h=h.ctypes.data, w=w.ctypes.data,
                                pdiag=pd.ctypes.data, enuclr=s.enuclr, atheat=s.atheat, scfcrt=scfcrt,
                                itrmax=itrmax)
        self.h = ctypes.c_void_p()
        L.check(self.lib.mhip_batch_create(len(systems), mols, ctypes.byref(self.h)))
        self._keep = None  # the library holds its own copies
        self.set_tiles(tiles or load_tiles())

    def set_tiles(self, tiles):
        t = (ctypes.c_int * 4)(*tiles)
        L.check(self.lib.mhip_batch_set_tiles(self.h, t))
        self.tiles = list(tiles)

    def set_eig_lds(self, on=True):
        """Eigensolver matrices in LDS (default) or in global memory."""
        L.check(self.lib.mhip_batch_set_eig_lds(self.h, 1 if on else 0))

    def eig_profile(self, iters):
        """Per size class: eigensolver launch time (ms, HIP events, eager
        launches), FP64 flops and applied rotations over `iters` iterations."""
        ms = np.zeros(4)
        fl = np.zeros(4)
        rot = np.zeros(4, dtype=np.int64)
        L.check(self.lib.mhip_batch_eig_profile(self.h, iters, L.ptr(ms), L.ptr(fl), L.ptr(rot)))
        return ms, fl, rot

    def run(self, iters_per_graph=40, max_launches=50):
        sec = ctypes.c_double()
        nconv = ctypes.c_int()
        L.check(self.lib.mhip_batch_run(self.h, iters_per_graph, max_launches, ctypes.byref(sec),
                                        ctypes.byref(nconv)))
        return sec.value, nconv.value

    def results(self):
        n = len(self.systems)
        e = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        cv = np.zeros(n, dtype=np.int32)
        L.check(self.lib.mhip_batch_results(self.h, L.ptr(e), L.ptr(it), L.ptr(cv)))
        return e, it, cv.astype(bool)

    def density(self, m):
        s = self.systems[m]
        p = np.zeros(s.mpack)
        L.check(self.lib.mhip_batch_density(self.h, m, L.ptr(p)))
        return p

    def close(self):
        if self.h:
            self.lib.mhip_batch_destroy(self.h)
            self.h = ctypes.c_void_p()


def build_systems(mols):
    return [hamiltonian.build(nat, xyz) for nat, xyz in mols]


def autotune(batch, repeats=2, write=True):
    """Coordinate search over the eigensolver workgroup size of each class."""
    best = list(batch.tiles)
    timings = {}
    batch.run()  # warm the caches and the graph path once
    for cls in range(4):
        scores = {}
        for cand in CANDIDATES:
            trial = list(best)
            trial[cls] = cand
            batch.set_tiles(trial)
            t = min(batch.run()[0] for _ in range(repeats))
            scores[cand] = t
            print(f"[tune] class {cls} tile {cand}: {t:.4f} s", flush=True)
        best[cls] = min(scores, key=scores.get)
        timings[f"class{cls}"] = scores
    batch.set_tiles(best)
    if write:
        os.makedirs(os.path.dirname(TILES_JSON), exist_ok=True)
        with open(TILES_JSON, "w") as fh:
            json.dump({"tiles": best, "classes": ["n<=32", "n<=64", "n<=128", "n<=256"],
                       "seconds_per_candidate": timings, "molecules": len(batch.systems),
                       "when": time.strftime("%Y-%m-%d %H:%M:%S")}, fh, indent=1)
    return best, timings


def main(argv=None):
    """python -m mopac_gpu_amd.batch [--out FILE]: tune the eigensolver
    workgroup sizes on the 4096-molecule workload (seed 4) and write the
    table (default: the package's tuning file)."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", default=TILES_JSON)
    ap.add_argument("--molecules", type=int, default=4096)
    ap.add_argument("--repeats", type=int, default=2)
    args = ap.parse_args(argv)
    systems = build_systems(workload(args.molecules, seed=4))
    b = Batch(systems, scfcrt=1e-6, tiles=[64, 128, 256, 256])
    try:
        best, timings = autotune(b, repeats=args.repeats, write=False)
        tuned = min(b.run()[0] for _ in range(args.repeats))
    finally:
        b.close()
    table = {"tiles": best, "classes": ["n<=32", "n<=64", "n<=128", "n<=256"],
             "seconds_per_candidate": timings, "molecules": len(systems), "tuned_seconds": tuned,
             "workload": f"batch.workload({len(systems)}, seed=4): alkanes CnH2n+2 (n=1..20) and water "
                         "clusters (1..30 H2O), SCFCRT=1e-6",
             "device": "MI355X (gfx950)"}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as fh:
        json.dump(table, fh, indent=1)
    print(json.dumps({"tiles": best, "tuned_seconds": tuned}), flush=True)


if __name__ == "__main__":
    main()
