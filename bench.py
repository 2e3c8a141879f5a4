#!/usr/bin/env python3
"""SCF time-to-converge on one MI355X: accelerated form vs plain FP64.

Workload: (H2O)512 water cluster (O on a 3.1 A 8x8x8 lattice, +-0.2 A jitter,
seed 2), PM7 RHF 1SCF with SCFCRT=1e-6: 3072 basis functions, 2048 doubly
occupied levels.  One "step" is one complete SCF from the starting guess to
self-consistency (guess + iterations + final diagonalization), with H and W
already resident in HBM.

  plain : the reference algorithm (full FP64 diagonalization every iteration)
  fast  : the accelerated form (learned starting density when a trained model
          is present, adaptive eigensolver tolerance, split-BF16 density while
          far from convergence), compared with plain FP64 by |dHf| and
          max|dP|.

Prints one JSON line (value = fast-form seconds per SCF, lower is better)
carrying the plain-FP64 time, the per-result error against FP64, a roofline
entry for the dominant kernel (timed with HIP events on the library stream)
and the CPU baseline (the FP64 restatement on the host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCF time-to-converge (s) at |ΔHf|<1e-3 kcal/mol vs FP64: mixed+learned vs plain"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
FAST_PRECISION = 1
HELIX_SCFCRT = 1.0e-6  # SCFCRT=1.D-6 for the MOZYME helix


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--waters", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--precision", type=int, default=FAST_PRECISION, choices=(0, 1),
                    help="fast form: 1 = split-BF16 contractions while far from convergence")
    ap.add_argument("--fast-only", action="store_true",
                    help="time only the fast form (no plain runs, no error report, no CPU leg): "
                         "the command the kernel profiles are taken from")
    ap.add_argument("--workload", choices=("water512", "helix"), default="water512",
                    help="helix: the MOZYME workload (poly-glycine alpha helix) instead of the "
                         "(H2O)512 metric line")
    ap.add_argument("--residues", type=int, default=400, help="helix length for --workload helix")
    ap.add_argument("--no-conventional", action="store_true",
                    help="helix: skip the conventional FP64 SCF of the same helix")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="instead of the SCF, copy a MIB-MiB buffer 10 times (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 8
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(10):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_read_per_launch": n * 8,
                      "bytes_written_per_launch": n * 8, "launches": 10}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the SCF of one molecule runs on a single "
                                   "device here, so --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.workload == "helix":
        run_helix(args)
        return
    rank = int(os.environ.get("RANK", "0"))
    import ctypes
    from mopac_gpu_amd import geometry, hamiltonian
    from mopac_gpu_amd import lib as L
    from mopac_gpu_amd import guess as learned
    from mopac_gpu_amd.scf import Context

    k = args.waters
    shape = (8, 8, 8) if k == 512 else None
    nat, xyz = geometry.water_cluster(k, 2 if k == 512 else 1, shape=shape)
    s = hamiltonian.build(nat, xyz)
    ctx = Context(0)
    dev_name, hbm, ncu = ctx.device_info()
    ctx.upload(s)
    lib = L.load()
    lib.mhip_set_tuning(ctx.h, 1, 1, 1.0e-10)

    model = learned.load_default()
    frag = geometry.fragments(nat, xyz)
    t_guess = 0.0

    def run_plain():
        return ctx.iter(scfcrt=1e-6, want=("p", "eig"))

    def run_fast(learned_guess=True, precision=None):
        nonlocal t_guess
        t0 = time.perf_counter()
        # learned starting density, built on the device (descriptors, network,
        # assembly); the standard guess when the system is not all water
        have = learned_guess and model is not None and model.to_device(ctx, s) is not None
        t_guess = time.perf_counter() - t0
        prec = args.precision if precision is None else precision
        r = ctx.iter(scfcrt=1e-6, adaptive_eigen=True, precision=prec, start_from_guess=have,
                     pseudo_diag=2, block_start=2, fragment=frag, want=("p", "eig"))
        r["seconds"] += t_guess
        return r

    # reference FP64 answer for the error report, and the plain-form timing
    ref = None
    plain_t = []
    if not args.fast_only:
        ref = run_plain()
        for _ in range(args.warmup):
            run_plain()
        for _ in range(args.steps):
            t0 = time.perf_counter()
            run_plain()
            plain_t.append(time.perf_counter() - t0)
    # accelerated-form timing (the metric)
    for _ in range(args.warmup):
        run_fast()
    fast_t = []
    fast_last = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        fast_last = run_fast()
        fast_t.append(time.perf_counter() - t0)
    t_fast = float(np.mean(fast_t))
    # the same accelerated form from the standard (pdiag) guess, for what the
    # learned starting density gains (not part of the metric)
    std_guess = None
    if not args.fast_only and model is not None:
        run_fast(learned_guess=False)
        t0 = time.perf_counter()
        r_std = run_fast(learned_guess=False)
        std_guess = {"seconds": round(time.perf_counter() - t0, 4), "iterations": r_std["niter"],
                     "dHf_kcal": abs(r_std["escf"] - ref["escf"]) if ref else None}
        run_fast()  # leave the context as the measured form left it
    # both precision settings of the fast form side by side (not the metric):
    # what the split-BF16 contractions gain and what they cost in error
    forms = None
    if not args.fast_only:
        forms = {}
        for prec in (0, 1):
            run_fast(precision=prec)
            t0 = time.perf_counter()
            r_p = run_fast(precision=prec)
            forms[str(prec)] = {"seconds": round(time.perf_counter() - t0, 4), "iterations": r_p["niter"],
                                "split_iterations": r_p["n_split_iter"],
                                "dHf_kcal": abs(r_p["escf"] - ref["escf"]),
                                "max_dP": float(np.max(np.abs(r_p["p"] - ref["p"])))}
        run_fast()
    t_plain = float(np.mean(plain_t)) if plain_t else None
    dhf = abs(fast_last["escf"] - ref["escf"]) if ref else None
    dp = float(np.max(np.abs(fast_last["p"] - ref["p"]))) if ref else None

    # ---- roofline: dominant kernel of the measured (fast) form, HIP events on
    # the stream every kernel is launched on
    n = s.norbs
    jv = [ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double(), ctypes.c_longlong(),
          ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong(),
          ctypes.c_longlong()]
    jrefs = [ctypes.byref(v) for v in jv]
    nk = 4
    kms, kl, kf, kb = (np.zeros(nk), np.zeros(nk, dtype=np.int64), np.zeros(nk), np.zeros(nk))
    lib.mhip_kernel_timing(ctx.h, 1, *jrefs)
    lib.mhip_kernel_stats(ctx.h, 1, nk, None, None, None, None)
    run_fast()
    lib.mhip_kernel_stats(ctx.h, -1, nk, L.ptr(kms), L.ptr(kl), L.ptr(kf), L.ptr(kb))
    lib.mhip_kernel_stats(ctx.h, 0, nk, None, None, None, None)
    lib.mhip_kernel_timing(ctx.h, 0, *jrefs)
    solve_ms, n_solve, apply_ms, n_apply, applyc_ms, n_applyc, tiles, pairs, rotations = [v.value for v in jv]
    npad = ((n + 63) // 64) * 64
    # block-Jacobi algorithmic work: A-update = two 64^3 products per processed
    # tile; C-update = one n x 64 x 64 product per processed pair; solve = the
    # rotations the library counted (32 x 32 cross rounds per pair, plus 2 x 496
    # inside the blocks once per sweep) x (64 + 64 + 64 element pairs) x 6 flops
    entries = [
        ("k_jac_solve", "mfma", solve_ms, n_solve, rotations * 3.0 * 64 * 6, 0.0),
        ("k_jac_apply_a_mfma", "mfma", apply_ms, n_apply, tiles * 2.0 * 2.0 * 64 ** 3, 0.0),
        ("k_jac_apply_c_mfma", "mfma", applyc_ms, n_applyc, pairs * 2.0 * npad * 64 * 64, 0.0),
    ]
    for i in range(nk):
        name = lib.mhip_kernel_class_name(i).decode()
        bound = "hbm" if "fock" in name else "mfma"
        entries.append((name, bound, float(kms[i]), int(kl[i]), float(kf[i]), float(kb[i])))
    def rate(e):
        if e[2] <= 0:
            return None
        return round((e[5] / 1e9 if e[1] == "hbm" else e[4] / 1e12) / (e[2] * 1e-3), 2)

    breakdown = {e[0]: {"ms": round(e[2], 3), "launches": int(e[3]),
                        ("GB/s" if e[1] == "hbm" else "TFLOP/s"): rate(e)} for e in entries if e[3] > 0}
    name, bound, tot_ms, nl, flops, nbytes = max(entries, key=lambda e: e[2])
    avg_s = (tot_ms / max(nl, 1)) * 1e-3
    if bound == "hbm":
        achieved = (nbytes / max(nl, 1)) / avg_s / 1e9 if avg_s > 0 else 0.0
        peak, unit = PEAK_HBM_GBS, "GB/s"
    else:
        achieved = (flops / max(nl, 1)) / avg_s / 1e12 if avg_s > 0 else 0.0
        peak = PEAK_FP64_TFLOPS
        unit = "TFLOP/s"
        if "split" in name:
            # the split-BF16 GEMM forms six BF16 products per FP64-equivalent
            # one: its matrix-core rate is 6x the FP64-equivalent rate, and it
            # is judged against the BF16 peak
            achieved *= 6.0
            flops *= 6.0
            peak = PEAK_BF16_TFLOPS
    # HBM bytes per launch of that kernel class, measured by rocprofv3 PMC runs
    # (FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size) of the same
    # workload; the numbers and the commands are in profiles/
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_h2o512_fast.json")
    if k == 512 and os.path.exists(tfile):
        with open(tfile) as fh:
            tc = json.load(fh).get("classes", {}).get(name)
        if tc:
            traffic = round(tc["traffic_bytes"])
    roof = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": round(nbytes / max(nl, 1)) if nbytes else None, "kernel": name,
            "avg_launch_us": round(avg_s * 1e6, 2), "launches": int(nl),
            "algorithmic_per_launch": round((nbytes if bound == "hbm" else flops) / max(nl, 1)),
            "kernel_ms_in_one_fast_scf": breakdown}

    # ---- CPU baseline: the FP64 restatement (MKL dsyevd) on host cores
    cpu = None
    if rank == 0 and not args.no_cpu and not args.fast_only:
        from oracle import oracle
        olib = oracle.load()
        # every core this process may run on (its CPU affinity: on the GPU box
        # the cores granted to the job), unless the thread count is set
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
        olib.o_set_threads(threads)
        t0 = time.perf_counter()
        o = oracle.iterate(s, scfcrt=1e-6)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(t_cpu, 3), "unit": "s", "cores": threads, "kind": "port",
               "sample": f"one complete SCF on the same (H2O){k} workload ({o['niter']} iterations, "
                         f"MKL dsyevd={bool(olib.o_lapack_available())}), same metric span",
               "escf_minus_gpu_plain_kcal": round(o["escf"] - ref["escf"], 8)}

    out = {
        "metric": METRIC,
        "value": round(t_fast, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_fast * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64 + bf16 (split-BF16 MFMA contractions while far from convergence; "
                  "BF16-weight network for the starting density)" if args.precision == 1 else
                  "f64 (BF16-weight network for the starting density)"),
        "data": "synthetic seeded (H2O)512 lattice cluster; PM7 integrals generated in-process",
        "config": {"workload": f"(H2O){k} PM7 RHF 1SCF SCFCRT=1e-6", "norbs": n, "nocc": s.nclose,
                   "form": ("mixed+learned" if model is not None else "mixed (no trained model found)"),
                   "precision": args.precision},
        "plain_fp64_s": round(t_plain, 4) if t_plain else None,
        "speedup_vs_plain": round(t_plain / t_fast, 3) if t_plain else None,
        "iterations": {"plain": ref["niter"] if ref else None, "fast": fast_last["niter"],
                       "plain_full_diag": ref["n_full_diag"] if ref else None,
                       "fast_full_diag": fast_last["n_full_diag"],
                       "plain_jacobi_sweeps": ref["jacobi_sweeps"] if ref else None,
                       "fast_jacobi_sweeps": fast_last["jacobi_sweeps"],
                       "fast_split_iterations": fast_last["n_split_iter"]},
        "error_vs_fp64": {"dHf_kcal": dhf, "max_dP": dp},
        "heat_of_formation_kcal": {"fp64": round(s.hof(ref["escf"]), 6) if ref else None,
                                   "fast": round(s.hof(fast_last["escf"]), 6),
                                   "dispersion": round(s.e_disp, 6), "hydrogen_bonds": round(s.e_hbond, 6)},
        "guess_s": round(t_guess, 4),
        "fast_from_standard_guess": std_guess,
        "precision_forms": forms,
        "device": dev_name,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def run_helix(args):
    """The MOZYME workload: a poly-glycine alpha helix (phi -57, psi -47,
    seeded +-5 degree torsion noise, H / OH caps), PM7 MOZYME 1SCF with
    SCFCRT=1.D-6 and MOZYME's default cutoffs (pairs stored within 9.9 A,
    dipole shell to 10 A, resonance to 7 A), FP64.  One step = the
    Lewis-structure start (hybrids, bonds polarized by the Fock matrix of the
    diagonal density) + the MOZYME SCF, with the sparse H and integrals
    resident.  The conventional FP64 SCF of the same helix on the GPU is run
    once beside it: the difference of the two heats of formation is MOZYME's
    cutoff approximation, not an error of either solver."""
    from mopac_gpu_amd import geometry, hamiltonian, mozyme
    from mopac_gpu_amd.scf import Context
    nres = args.residues
    nat, xyz = geometry.glycine_helix(nres, 3)
    t0 = time.perf_counter()
    ms = mozyme.MozymeSystem.from_geometry(nat, xyz)
    t_setup = time.perf_counter() - t0
    progress(f"helix Gly{nres}: {ms.numat} atoms, {ms.norbs} orbitals, mpack {ms.mpack}, setup {t_setup:.2f} s")
    ctx = Context(0)
    dev_name, _, _ = ctx.device_info()
    ms.upload(ctx)

    def step():
        t = time.perf_counter()
        f0 = ms.fock2z(ctx, ms.diagonal_density())
        lm = mozyme.lewis_lmos(ms, f0)
        t_start = time.perf_counter() - t
        r = ms.scf(ctx, lm, scfcrt=HELIX_SCFCRT)
        return time.perf_counter() - t, t_start, r

    for _ in range(args.warmup):
        dt, _, r = step()
        progress(f"warmup step {dt:.3f} s, {r['niter']} iterations")
    times, last = [], None
    for _ in range(args.steps):
        dt, t_start, last = step()
        times.append(dt)
        progress(f"timed step {dt:.3f} s, {last['niter']} iterations, escf {last['escf']:.6f}")
    t_mz = float(np.mean(times))
    # roofline of the dominant kernel, the diagg2 rotation levels: HIP events
    # around the level launches, algorithmic bytes counted by the library
    import ctypes
    from mopac_gpu_amd import lib as L
    lib = L.load()
    sv = [ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()]
    lib.mhip_mozyme_stats(1, *[ctypes.byref(v) for v in sv])
    step()
    lib.mhip_mozyme_stats(0, *[ctypes.byref(v) for v in sv])
    d2_ms, d2_rot, d2_lev, d2_launch, d2_bytes = [v.value for v in sv]
    avg_s = d2_ms * 1e-3 / max(d2_launch, 1)
    achieved = (d2_bytes / max(d2_launch, 1)) / avg_s / 1e9 if avg_s > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": None, "kernel": "k_mz_diagg2",
            "avg_launch_us": round(avg_s * 1e6, 2), "launches": int(d2_launch),
            "algorithmic_bytes_per_launch": round(d2_bytes / max(d2_launch, 1)),
            "rotations_per_scf": int(d2_rot), "levels_per_scf": int(d2_lev), "ms_per_scf": round(d2_ms, 2),
            "note": "one level per launch; each rotation is a chain of dependent gathers, so the kernel is "
                    "latency-bound far below the HBM roof"}
    # CPU baseline: the reference's own MOZYME routines (compiled from its
    # sources into oracle/_ref), one core, a bounded sample of iterations
    cpu = None
    if not args.no_cpu:
        cpu = helix_cpu_baseline(ms, last["niter"])
    conv = None
    if not args.no_conventional:
        progress("conventional SCF of the same helix")
        s = hamiltonian.build(nat, xyz)
        c2 = Context(0)
        c2.upload(s)
        rc = c2.iter(scfcrt=1e-6)
        conv = {"seconds": round(rc["seconds"], 3), "iterations": rc["niter"], "escf": round(rc["escf"], 6),
                "mozyme_minus_conventional_kcal": round(last["escf"] - rc["escf"], 6)}
        c2.close()
    out = {
        "metric": "MOZYME SCF time-to-converge (s), poly-glycine alpha helix",
        "value": round(t_mz, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_mz * 1e3, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": f"synthetic seeded Gly{nres} alpha helix; PM7 integrals generated in-process",
        "config": {"workload": f"Gly{nres} alpha helix PM7 MOZYME 1SCF SCFCRT=1.D-6", "numat": ms.numat,
                   "norbs": ms.norbs, "nocc": int(ms.s.nclose), "mpack": int(ms.mpack),
                   "stored_pairs_within_A": mozyme.NDDO_CUTOFF, "dipole_shell_A": mozyme.DIPOLE_CUTOFF,
                   "resonance_within_A": ms.overlap_cutoff, "scfcrt": HELIX_SCFCRT},
        "iterations": last["niter"], "converged": last["converged"], "escf": round(last["escf"], 6),
        "lewis_start_s": round(t_start, 4), "sparse_setup_s": round(t_setup, 3),
        "conventional_fp64": conv, "device": dev_name, "roofline": roof, "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


def helix_cpu_baseline(ms, niter, sample_iters=6):
    """One core running the reference's density_for_MOZYME, fock2z (fz2n),
    helecz, diagg1 and diagg2 (oracle/_ref) from the same Lewis start, for
    `sample_iters` iterations; the per-iteration time is scaled to the GPU
    SCF's iteration count and the start added."""
    from oracle import ref as REF
    from mopac_gpu_amd import mozyme
    if not REF.available():
        return None
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    t0 = time.perf_counter()
    p0 = ms.diagonal_density()
    f0 = REF.mozyme_fock2z(ms, p0, feather=True)
    lm = mozyme.lewis_lmos(ms, f0)
    t_start = time.perf_counter() - t0
    lm.fao, lm.p = f0, p0
    lm.eigs, lm.eigv = np.zeros(lm.nocc), np.zeros(lm.nvir)
    R = REF.Mozyme(mozyme.reference_capacity(lm))
    lib, P = REF.load(), REF.P
    xyz, am, dd, ad, w, h = f64(ms.xyz), f64(ms.am), f64(ms.dd), f64(ms.ad), f64(ms.w), f64(ms.h)
    t1 = time.perf_counter()
    for it in range(1, sample_iters + 1):
        progress(f"cpu baseline iteration {it}")
        p = R.density(0)
        f = np.zeros(ms.mpack)
        lib.mz_fock2z(P(xyz), P(am), P(dd), P(ad), len(w), P(w), P(p), P(h), REF.A0, REF.EV, 1, REF.TRUNC1,
                      REF.TRUNC2, P(f))
        R.helecz(p, h, f)
        r = R.diagg1(f, it)
        R.diagg2(r["nij"], it)
    per_iter = (time.perf_counter() - t1) / sample_iters
    return {"value": round(t_start + per_iter * niter, 3), "unit": "s", "cores": 1, "kind": "reference",
            "sample": f"{sample_iters} iterations of the reference's density_for_MOZYME, fock2z, helecz, diagg1 "
                      f"and diagg2 on the same helix from the same Lewis start ({per_iter:.3f} s each), scaled "
                      f"to the GPU SCF's {niter} iterations, plus the start ({t_start:.2f} s)"}


if __name__ == "__main__":
    main()
