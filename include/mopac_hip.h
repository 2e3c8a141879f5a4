/*
 * mopac_hip.h — C ABI of libmopachip, a gfx950 (MI355X) implementation of
 * MOPAC's SCF inner loop.
 *
 * Conventions (the reference's Fortran conventions, so a Fortran host can bind
 * every entry point with ISO_C_BINDING):
 *   - dense matrices are column-major with leading dimension n;
 *   - symmetric matrices are MOPAC "packed": the lower triangle by rows,
 *     element (i,j), i>=j, 1-based, at i(i-1)/2 + j (mpack = n(n+1)/2);
 *   - orbital ranges nfirst/nlast are 1-based;
 *   - every function returns 0 on success and a nonzero status on error;
 *     mhip_last_error() returns a message for the last failure on this thread.
 *   - pointers are host memory unless the name ends in _dev.
 *
 * One GPU per context.  A context is not thread-safe (like the reference's
 * global handles); distinct contexts may be driven from distinct threads.
 *
 * Citations name the reference interface each entry point replaces
 * (paths relative to the reference repository root).
 */
#ifndef MOPAC_HIP_H
#define MOPAC_HIP_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhip_ctx mhip_ctx;

/* ---- lifecycle --------------------------------------------------------- */

/* Open device `device` and create a context.
 * Replaces getGPUInfo / setDevice / create_handle
 * (src/gpu/cuda_wrappers.cu:71-118, :158-187). */
int mhip_init(int device, mhip_ctx** ctx);

/* Release every device buffer of the context.  Idempotent for NULL.
 * Replaces mopac_cuda_destroy_resources (src/gpu/cuda_wrappers.cu:830-856). */
int mhip_finalize(mhip_ctx* ctx);

const char* mhip_last_error(void);

/* Device name and HBM size, for logs (getGPUInfo, cuda_wrappers.cu:71-112). */
int mhip_device_info(mhip_ctx* ctx, char* name, int name_len, double* hbm_gib, int* n_cu);

/* ---- system upload (once per geometry) --------------------------------- */

/* Make H (mpack) and W (n2elec, hcore order) resident on the device and build
 * the atom-pair tables the Fock kernels walk.  Data produced by hcore
 * (src/integrals/hcore.F90:157-288) crosses PCIe once, not once per call. */
int mhip_upload_system(mhip_ctx* ctx, int numat, const int* nfirst, const int* nlast, int norbs,
                       long long n2elec, const double* h, const double* w);

/* ---- the hot path, one routine at a time -------------------------------- */

/* P = occ * C(:,1:ndubl) * C(:,1:ndubl)^T, packed into p (mpack).
 * c == NULL uses the context's resident eigenvectors; p == NULL leaves the
 * result resident.  Semantics of density_for_GPU case 4/5 for fract < 1e-2
 * (src/matrix/density_for_GPU.F90:92-108); replaces call_syrk_cublas
 * (src/gpu/cuda_wrappers.cu:229-254) on that path. */
int mhip_density(mhip_ctx* ctx, const double* c, int ndubl, double occ, double* p);

/* F += G(P): two-centre Coulomb/exchange over atom pairs plus (mode == 2) the
 * one-centre terms.  f holds H (or the shifted H) on entry.  ptot = total
 * density, p = alpha density.  Semantics of fock2(f,ptot,p,w,w,w,numat,
 * nfirst,nlast,mode) for molecules (src/SCF/fock2.F90:16-374, fock1.F90). */
int mhip_fock2(mhip_ctx* ctx, double* f, const double* ptot, const double* p, int mode);

/* Electronic energy E = sum_{i>j} P(H+F) + 1/2 sum_i P(H+F)
 * (src/integrals/helect.F90:16-66). */
int mhip_helect(mhip_ctx* ctx, int norbs, const double* p, const double* h, const double* f,
                double* e);

/* Allocation-free variant of mhip_helect.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc).  p, h, f
 * (packed, mpack each) and e (one double) are DEVICE pointers owned by the
 * caller; the work is queued on `stream` (a hipStream_t, NULL = the default
 * stream) and is not waited for.  The per-row partial sums live in a fixed
 * device array of mhip_helect_noalloc_max_rows() doubles per GPU (65 536,
 * 512 KiB, held until the process ends); larger norbs is rejected with an
 * error.  Calls no hipMalloc/hipFree.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (different streams or threads)
 * race.  Returns 0 on success. */
int mhip_helect_noalloc(int norbs, const double* p, const double* h, const double* f, double* e,
                        void* stream);
int mhip_helect_noalloc_max_rows(void);

/* `ncalls` Pulay DIIS steps from a fresh start on the uploaded system, as
 * iter makes them (src/SCF/pulay.F90:16-166, 10-deep circular history,
 * osinv): call k extrapolates f_seq[k] (mpack) in place from the alpha
 * density pa_seq[k], and pl_seq[k] receives its PL.  The same device step
 * mhip_scf takes with PULAY, exposed so a sequence that wraps the history can
 * be checked call for call. */
int mhip_pulay_sequence(mhip_ctx* ctx, int ncalls, double* f_seq, const double* pa_seq, double* pl_seq);

/* All eigenpairs of the packed Fock matrix.  Adds i*1e-10 to F(i,i) IN PLACE
 * (as the reference does), returns eigenvectors as columns of c (n x n) in
 * ascending eigenvalue order.  Semantics of eigenvectors_LAPACK(c,f,eig,n)
 * (src/matrix/eigenvectors_LAPACK.F90:16-127); replaces mopac_cuda_dsyevd
 * (src/gpu/cuda_wrappers.cu:469-505). */
int mhip_eigen(mhip_ctx* ctx, double* c, double* fpacked, double* eig, int ndim);

/* Pseudo-diagonalization: FMO = C_virt^T F C_occ, then the ordered 2x2
 * occupied-virtual rotations of `vector`.  Semantics of
 * diag(fao,vector,nocc,eig,mdim,n) (src/matrix/diag.F90:16-163); replaces
 * call_rot_cuda_gpu (src/gpu/cuda_wrappers.cu:547-623).  *nrot receives the
 * number of rotations applied (may be NULL). */
int mhip_pseudo_diag(mhip_ctx* ctx, const double* fao, double* vector, int nocc,
                     const double* eig, int mdim, int n, long long* nrot);

/* ---- the whole SCF, resident on the device ------------------------------ */

typedef struct {
  int nclose;          /* doubly occupied levels */
  int pulay;           /* PULAY keyword (iter.F90:131) */
  int itrmax;          /* ITRY= (default 2000) */
  double scfcrt;       /* SCFCRT= (default 1e-4) */
  int last;            /* 1 for a 1SCF job (phase_lock at exit) */
  int precision;       /* 0 = FP64; 1 = split-BF16 MFMA contractions (density, block
                          pseudo-diagonalization) while PL > 100 x PLTEST, FP64 after */
  int pseudo_diag;     /* once newdg holds (iter.F90:843-861): 0 = always full diagonalization
                          (the reference), 1 = diag.F90's ordered rotation list, 2 = the same
                          rotations applied at once as GEMMs + Bjorck re-orthonormalization */
  int adaptive_eigen;  /* 1 = loosen the eigensolver tolerance while PL is large */
  double enuclr;       /* eV */
  double atheat;       /* kcal/mol */
  const double* pdiag; /* starting diagonal (moldat.F90:731-790), n */
  const double* p_init;/* optional starting total density (mpack), overrides pdiag */
  int block_start;     /* with pseudo_diag == 2: 0 = block steps once newdg holds; 1 = whenever
                          eigenvectors exist and came from a supplied starting density or the
                          fragment blocks (from the standard diagonal guess alone, once newdg
                          holds); 2 = also take the first eigenvectors from the fragment
                          blocks of F (needs fragment and pdiag) */
  const int* fragment; /* optional atom -> fragment id (0-based, e.g. the molecules of a
                          cluster; <= 16 orbitals and an even electron count each) */
  int start_from_guess;/* 1 = start from the density mhip_guess_water left on the device
                          (overrides p_init and pdiag) */
} mhip_scf_opts;

typedef struct {
  double escf;         /* heat of formation, kcal/mol (without geometry-only terms) */
  double ee;           /* electronic energy, eV */
  int niter;
  int converged;
  int n_full_diag;
  int n_pseudo_diag;
  double seconds;      /* wall time of guess + iterations + final diag */
  int jacobi_sweeps;   /* block-Jacobi sweeps summed over all diagonalizations */
  int n_fragment_diag; /* fragment-block starts (0 or 1) */
  int n_split_iter;    /* iterations whose contractions ran split-BF16 */
} mhip_scf_result;

/* Run iter() (src/SCF/iter.F90:16-1107, RHF closed shell) with every matrix
 * resident in HBM.  Outputs (any may be NULL): p (mpack total density),
 * f (mpack final Fock), c (n x n eigenvectors), eig (n). */
int mhip_scf(mhip_ctx* ctx, const mhip_scf_opts* opts, mhip_scf_result* res, double* p,
             double* f, double* c, double* eig);

/* ---- batched SCF: many small molecules in one captured graph --------------
 * No single counterpart in the reference (it runs one molecule per job); each
 * molecule follows iter (src/SCF/iter.F90:16-1107, RHF, no PULAY) with its
 * control logic on the device.  Molecules up to 256 basis functions; the
 * arrays of each molecule are the ones mhip_upload_system / mhip_scf take. */
typedef struct {
  int numat, norbs, nclose;
  const int* nfirst;       /* 1-based */
  const int* nlast;
  long long n2elec;
  const double* h;         /* packed */
  const double* w;         /* hcore order */
  const double* pdiag;     /* starting diagonal */
  double enuclr, atheat;   /* eV, kcal/mol */
  double scfcrt;           /* SCFCRT (default 1e-4 when <= 0) */
  int itrmax;              /* default 2000 when <= 0 */
} mhip_batch_molecule;

typedef struct mhip_batch mhip_batch;
int mhip_batch_create(int nmol, const mhip_batch_molecule* mols, mhip_batch** out);
/* Eigensolver workgroup size per size class (n <= 32, 64, 128, 256):
 * 64, 128, 256, 512 or 1024 (chosen by the autotuner, mopac_gpu_amd/batch.py). */
int mhip_batch_set_tiles(mhip_batch* b, const int* tiles);
/* Where the eigensolver keeps its matrices: lds = 1 (default) puts the
 * packed A = C^T F C, and C where both fit, in LDS; lds = 0 keeps both in
 * global memory (the same rotations, for comparison). */
int mhip_batch_set_eig_lds(mhip_batch* b, int lds);
/* Eigensolver profile: `iters` iterations of the batch from its starting
 * densities, launched one by one (no graph) with HIP events around the
 * eigensolver launch of each size class (n <= 32, 64, 128, 256).  ms[4]: the
 * classes' summed launch times; flops[4]: their FP64 flops (12 n per applied
 * rotation, 2 n^3 + n^2 (n+1) for a warm start, n (n+1) nocc for the
 * density); rotations[4] (may be NULL): applied rotations. */
int mhip_batch_eig_profile(mhip_batch* b, int iters, double* ms, double* flops, long long* rotations);
/* Capture iters_per_graph iterations of the whole batch into one hipGraph
 * and replay it (at most max_launches times) until every molecule is done. */
int mhip_batch_run(mhip_batch* b, int iters_per_graph, int max_launches, double* seconds,
                   int* n_converged);
/* Heats of formation (kcal/mol, without geometry-only terms), iterations and
 * convergence flags of every molecule; any pointer may be NULL. */
int mhip_batch_results(mhip_batch* b, double* escf, int* niter, int* converged);
/* Packed total density of molecule m. */
int mhip_batch_density(mhip_batch* b, int m, double* p);
void mhip_batch_destroy(mhip_batch* b);

/* ---- MOZYME localized-MO kernels -------------------------------------------
 * The MOZYME_C module arrays, passed explicitly (C cannot see Fortran
 * modules), with the reference's conventions: atom numbers 1-based, offsets
 * 0-based (MOZYME_C.F90:116-162); nijbo numat x numat column-major, -1 where
 * an atom pair is not stored; the sparse P and F of length mpack hold the
 * atom-pair blocks at nijbo.  diagg2 grows LMOs inside the room each owns up
 * to the next LMO's offset (or icocc_dim / cocc_dim for the last), and
 * updates ncf/icocc/cocc and nce/icvir/cvir in place. */
typedef struct {
  int numat, norbs, nocc, nvir;
  long long mpack;
  const int* nfirst;   /* numat, 1-based */
  const int* nlast;    /* numat, 1-based */
  const int* nijbo;    /* numat * numat */
  int* ncf;            /* nocc */
  const int* nncf;     /* nocc */
  const int* ncocc;    /* nocc */
  int* icocc;          /* icocc_dim */
  int icocc_dim;
  double* cocc;        /* cocc_dim */
  int cocc_dim;
  int* nce;            /* nvir */
  const int* nnce;     /* nvir */
  const int* ncvir;    /* nvir */
  int* icvir;          /* icvir_dim */
  int icvir_dim;
  double* cvir;        /* cvir_dim */
  int cvir_dim;
} mhip_lmo_set;

/* density_for_MOZYME(p, mode, nclose, partp) over all nocc LMOs
 * (src/MOZYME/density_for_MOZYME.F90:16-342): mode 0 full density, -1 and 1
 * the partial forms with partp. */
int mhip_mozyme_density(mhip_ctx* ctx, const mhip_lmo_set* lmo, int mode, const double* partp,
                        double* p);

/* The values diagg1 keeps between calls (SAVE, diagg1.F90:61-71) and the
 * MOZYME_C scalars it reads and writes.  Set reset = 1 for a new SCF (oldens
 * as the OLDENS keyword); nfmo (nvir ints, caller-owned) carries each
 * virtual LMO's list length to the calls that reuse the list.
 * fp32_accumulate (no counterpart in the reference): 1 forms the
 * occupied-virtual elements (the F c products of each virtual LMO and their
 * dot products with the occupied LMOs) in FP32 from the FP64 data, the
 * mixed-precision form; the virtual LMO energies follow from those products,
 * while the occupied LMO energies, the tests' weights and everything after
 * diagg1 stay FP64. */
typedef struct {
  int reset, oldens;
  double fref, safety, oldlim;
  double tiny, sumt, ovmax;
  int* nfmo;
  int fp32_accumulate;
} mhip_diagg1_state;

/* diagg1(fao, nocc, nvir, eigv, ...) (src/MOZYME/diagg1.F90:16-649): the
 * virtual LMO energies eigv, the occupied ones eigs (on the iterations the
 * reference evaluates them) and the list of occupied-virtual elements
 * fmo(1:nij), ifmo(2, 1:nij) (virtual index first, 1-based), at most fmo_dim. */
int mhip_mozyme_diagg1(mhip_ctx* ctx, const mhip_lmo_set* lmo, const double* fao, const double* p,
                       double* eigs, double* eigv, int idiagg, mhip_diagg1_state* st, int fmo_dim,
                       double* fmo, int* ifmo, int* nij);

/* diagg2's saved rejection counts (diagg2.F90:55,338-339) and the MOZYME_C
 * scalars it uses: tiny (from diagg1, updated here), sumb (out), damp (the
 * DAMP constant, 1 for main-group molecules), shift, thresh (1e-13 by
 * default, iter_for_MOZYME.F90:86-97).  order_free (no counterpart in the
 * reference): 0 applies the rotations in list order as diagg2 does; 1 applies
 * the same rotations in rounds of LMO-disjoint pairs (a first-fit colouring
 * of the list), which changes the order of rotations that share an LMO and
 * shortens the chain of dependent rotations to about the largest number of
 * rotations on one LMO. */
typedef struct {
  int reset;
  double tiny, sumb, damp, shift, thresh;
  int nrejct[2];
  int order_free;
} mhip_diagg2_state;

/* diagg2(nocc, nvir, eigv, ..., nij, idiagg, ...) (src/MOZYME/diagg2.F90:
 * 16-359): the 2x2 annihilations of the listed elements, in list order, with
 * LMO growth, rejection and the half-angle retry. */
int mhip_mozyme_diagg2(mhip_ctx* ctx, mhip_lmo_set* lmo, const double* eigs, const double* eigv,
                       const double* fmo, const int* ifmo, int nij, int idiagg,
                       mhip_diagg2_state* st);
/* tidy (src/MOZYME/tidy.F90:16-244) at a single point on the occupied
 * (occupied = 1) or the virtual set of lmo, on the device: atoms whose
 * squared coefficient sum is <= thresh leave their LMO, and the free space
 * is spread over the LMOs as the reference spreads it (LMO 1 stays at 0).
 * Updates the set's counts, offsets, atom lists and coefficients in place
 * (the free space between LMOs is left undefined); fails when the storage
 * is too small for the next annihilation step. */
int mhip_mozyme_tidy(mhip_ctx* ctx, mhip_lmo_set* lmo, int occupied, double thresh);
/* check (src/MOZYME/check.F90:16-113) on the occupied or the virtual set, on
 * the device: every LMO scaled to unit norm; *error = the summed |1 - norm|
 * before the scaling (the reference stops the job above 0.1). */
int mhip_mozyme_check(mhip_ctx* ctx, mhip_lmo_set* lmo, int occupied, double* error);
/* eimp (src/MOZYME/eimp.F90:16-45) on the device: the first element of each
 * stored off-diagonal block of p becomes the sum of squares of that block of
 * f. */
int mhip_mozyme_eimp(mhip_ctx* ctx, int numat, const int* nfirst, const int* nlast, const int* nijbo,
                     long long mpack, const double* f, double* p);
/* No counterpart in the reference.  HIP-event time of the diagg2 rotation
 * kernels on the MOZYME stream, with the rotations, levels, launches,
 * algorithmic bytes (each rotation reads and writes both LMOs' coefficients,
 * sized on entry) and, for the one-workgroup kernel, the rounds (a level of
 * w rotations takes ceil(w / waves) rounds) since the last reset; enable =
 * 1 / 0 resets and switches the count on / off, -1 only reads. */
int mhip_mozyme_stats(int enable, double* ms, long long* rotations, long long* levels, long long* launches,
                      double* bytes, long long* rounds);
/* No counterpart in the reference.  Selects how mhip_mozyme_diagg2 runs the
 * rotation levels: 1 (default) on the LDS rotation kernels, a wave per
 * rotation with its atom maps and lists in LDS: runs of narrow levels in one
 * launch of a single workgroup that meets at a barrier between levels, and
 * each level wider than mhip_mozyme_diagg2_wide_rounds rounds of that
 * workgroup across the chip, a workgroup per rotation; they apply when
 * numat and norbs are below 65535 and the maps fit in LDS (two LMOs with
 * more than mhip_mozyme_diagg2_list_cap atoms keep their lists in global
 * memory).  0, or when they do not apply: one
 * launch per level of the kernel with global maps.  All give identical
 * results.  enable = -1 only reads; returns the previous setting. */
int mhip_mozyme_diagg2_one_workgroup(int enable);
/* No counterpart in the reference.  The width, in rounds of the one-workgroup
 * kernel, above which a level runs across the chip (default 1); rounds = -1
 * only reads; returns the previous value. */
int mhip_mozyme_diagg2_wide_rounds(int rounds);
/* No counterpart in the reference.  The LDS rotation kernels copy the two
 * LMOs' atom lists into LDS when they hold at most cap atoms together
 * (default and maximum 1024) and read them from global memory otherwise;
 * results are identical.  cap = -1 only reads; returns the previous value. */
int mhip_mozyme_diagg2_list_cap(int cap);
/* No counterpart in the reference.  The LDS rotation kernels copy the first
 * cap coefficients of each of the two LMOs into LDS (default and maximum
 * 1024) and, when every position a rotation can touch lies in those copies,
 * rotate from them; otherwise they read the coefficients from global memory
 * with an undo log.  Results are identical.  cap = -1 only reads; returns the
 * previous value. */
int mhip_mozyme_diagg2_coef_cap(int cap);

/* helecz() (src/MOZYME/helecz.F90:16-52): E = sum over stored atom pairs of
 * P (H + F), diagonal elements weighted 1/2, on the sparse nijbo storage. */
int mhip_mozyme_helecz(mhip_ctx* ctx, int numat, const int* nfirst, const int* nlast, const int* nijbo,
                       long long mpack, const double* p, const double* h, const double* f, double* e);

/* cnvgz(pnew, p, p1, p2, p3, niter, idiag) (src/MOZYME/cnvgz.F90:16-85):
 * three-point extrapolation every third cycle and diagonal damping from the
 * fourth (three_point = use_three_point_extrap); idiag(i) is the 1-based
 * position of P(i,i) in the sparse storage; pmax = max|pnew - p|.  All arrays
 * are updated in place as the reference updates them. */
int mhip_mozyme_cnvgz(mhip_ctx* ctx, int norbs, long long mpack, double* pnew, double* p, double* p1,
                      double* p2, double* p3, int niter, const int* idiag, int three_point, double* pmax);

/* The MOZYME system for the sparse Fock build: the one-electron matrix on the
 * sparse storage and the two-electron integrals as hcore_for_MOZYME stores
 * them for SEMIDIRECT (src/MOZYME/hcore_for_MOZYME.F90:100-405): for each atom
 * i, the blocks of the stored pairs (i, j < i) in order (100 / 10 / 10 / 1
 * values), then i's one-centre block; pairs that are not stored (nijbo -1,
 * or -2 for the dipole shell) take point-charge terms from the per-atom
 * parameters am, dd, ad (parameters_C), evaluated on the fly (semidr). */
typedef struct {
  int numat, norbs;
  long long mpack;      /* sparse storage length */
  const int* nfirst;    /* numat, 1-based */
  const int* nlast;
  const int* nijbo;     /* numat * numat, column-major, symmetric */
  const double* xyz;    /* 3 * numat, Angstrom */
  const double* am;     /* per atom */
  const double* dd;
  const double* ad;
  const double* h;      /* mpack */
  const double* w;      /* nw */
  long long nw;
  int feather;          /* l_feather (1 for PM7): smooth switch to point charges */
  double trunc1, trunc2;/* its radius (Angstrom) and Gaussian exponent (molkst_C trunc_1, trunc_2) */
} mhip_mozyme_system;
/* Upload a MOZYME system into the context (replaces the conventional one). */
int mhip_mozyme_setup(mhip_ctx* ctx, const mhip_mozyme_system* sys);
/* buildf(f, partf, 0) (src/MOZYME/buildf.F90:16-52): f = h + G(p) on the
 * sparse storage, G as fock2z/fz2n (src/MOZYME/fock2z.F90:16-1362) forms it for
 * a discrete system (ione = 1, direct = false, semidr = true). */
int mhip_mozyme_fock2z(mhip_ctx* ctx, const double* p, double* f);

/* The MOZYME SCF (iter_for_MOZYME, src/MOZYME/iter_for_MOZYME.F90:16-691)
 * for a single point of a discrete system, after mhip_mozyme_setup: starting
 * from the LMOs in lmo (updated in place, offsets included; their capacities
 * must leave room to grow), each iteration runs tidy, check, eimp, diagg
 * (diagg1, diagg2, density_for_MOZYME), cnvgz while the three-point
 * extrapolation is on, buildf/fock2z and helecz, with isitsc's test and the
 * reference's level-shift schedule, all on the device with the LMOs, P and F
 * resident.  The hydrogen-bond search addhb is not run. */
typedef struct {
  double scfcrt;    /* SELCON, kcal/mol (scfcri: 1e-2 when <= 0) */
  int itrmax;       /* 2000 when <= 0 */
  double thresh;    /* LMO coefficient threshold (1e-13 when <= 0) */
  double damp;      /* DAMP (1 when <= 0) */
  int fmo_dim;      /* occupied-virtual list capacity (0: automatic) */
  double enuclr;    /* eV */
  double atheat;    /* kcal/mol */
  int print_level;  /* 1: one line per iteration on stderr (iteration,
                       heat, its change, largest occupied-virtual element) */
  double* heats;    /* if not NULL: heat after iteration k at heats[k - 1], */
  int heats_cap;    /* for k <= heats_cap */
  int order_free;   /* diagg2's rotations in LMO-disjoint rounds
                       (mhip_diagg2_state.order_free); 0: list order */
  int fp32_fmo;     /* diagg1's occupied-virtual elements in FP32
                       (mhip_diagg1_state.fp32_accumulate) */
} mhip_mozyme_opts;
typedef struct {
  double escf;      /* heat of formation, kcal/mol (no geometry-only terms) */
  double ee;        /* electronic energy, eV */
  int niter;
  int converged;
  double ovmax;     /* largest occupied-virtual element of the last diagg1 */
  double seconds;
} mhip_mozyme_result;
int mhip_mozyme_scf(mhip_ctx* ctx, mhip_lmo_set* lmo, const mhip_mozyme_opts* opts, mhip_mozyme_result* res,
                    double* p, double* f);

/* ---- reference-boundary drop-ins -----------------------------------------
 * Same C signatures as the reference's exported symbols, so its Fortran
 * interfaces (src/gpu/cublas_interfaces.F90, src/matrix/eigenvectors_cuda.F90,
 * src/deprecated/mod_gpu_info.F90) bind by changing only the
 * bind(C, name=...) strings.  They use an implicit context on the device
 * mhip_setDevice selected (device 0 by default). */

/* call_gemm_cublas (src/gpu/cuda_wrappers.cu:192-226) */
void mhip_call_gemm(char tra, char trb, int m, int n, int k, double alpha, const double* a,
                    int lda, const double* b, int ldb, double beta, double* c, int ldc);
/* call_syrk_cublas (src/gpu/cuda_wrappers.cu:229-254); uplo 'U' or 'L' */
void mhip_call_syrk(char uplo, char tra, int n, int k, double alpha, const double* a, int lda,
                    double beta, double* c, int ldc);
/* mopac_cuda_dsyevd (src/gpu/cuda_wrappers.cu:469-505): A (upper) overwritten
 * by eigenvectors, W eigenvalues ascending, *info = 0 on success. */
void mhip_dsyevd(int n, double* a, int lda, double* w, int* info);
/* call_rot_cuda_gpu (src/gpu/cuda_wrappers.cu:547-623): fmo is the
 * (n-nocc) x nocc occupied-virtual block, virtual index slow. */
void mhip_call_rot(const double* fmo, const double* eig, double* v, const double* ci0,
                   const double* ca0, int nocc, int lumo, int n, double bigeps, double tiny);
/* The precision form's contraction with call_gemm's signature: FP64 in and
 * out, computed from a three-way BF16 split of each operand on the BF16
 * matrix cores (FP32 accumulation in 32-term chunks, summed in FP64).  No
 * counterpart in the reference; exported so the contraction's error against
 * FP64 can be measured on its own. */
void mhip_gemm_split_bf16(char tra, char trb, int m, int n, int k, double alpha, const double* a,
                          int lda, const double* b, int ldb, double beta, double* c, int ldc);
/* mopac_cuda_destroy_resources (src/gpu/cuda_wrappers.cu:830-856).
 * Idempotent; the next drop-in call sets the implicit context up again. */
void mhip_destroy_resources(void);
/* getGPUInfo (src/gpu/cuda_wrappers.cu:71-112), bound by
 * src/deprecated/mod_gpu_info.F90:5-15: up to six devices' names (the
 * marketing name, or the gfx target when the driver reports none), memory,
 * clock (kHz) and gfx major/minor; hasDouble is true for every device. */
void mhip_getGPUInfo(bool* hasGpu, bool hasDouble[6], int* nDevices, char name[6][256], int name_size[6],
                     size_t totalMem[6], int clockRate[6], int major[6], int minor[6]);
/* setDevice (src/gpu/cuda_wrappers.cu:115-118), bound by
 * mod_gpu_info.F90:24-28: selects the device the drop-ins above run on;
 * *stat false for an index that does not exist. */
void mhip_setDevice(int idevice, bool* stat);
/* set_mozyme_gpu_pair (src/gpu/cuda_wrappers.cu:58-68): the reference's
 * two-GPU MOZYME device pair.  Recorded (clamped to existing devices) for
 * link compatibility; this library runs MOZYME on one GPU. */
void mhip_set_mozyme_gpu_pair(int dev0, int dev1);
/* The device mhip_setDevice last selected (0 by default). */
int mhip_selected_device(void);

/* ---- learned starting density: the device MLP ------------------------------
 * No counterpart in the reference (it replaces nothing; it feeds the SCF a
 * starting density through mhip_scf_opts.p_init, as OLDENS feeds iter).
 * Row-major activations; layer l has dims[l] inputs and dims[l+1] outputs;
 * the flat parameter vector holds, per layer, W (dims[l] x dims[l+1],
 * row-major) then b.  BF16 weights and layer inputs, FP32 accumulation, tanh
 * hidden activations, linear output, mean-squared-error loss, Adam. */
typedef struct mhip_mlp mhip_mlp;
int mhip_mlp_create(int nlayers, const int* dims, mhip_mlp** out);
void mhip_mlp_destroy(mhip_mlp* mlp);
int mhip_mlp_set_params(mhip_mlp* mlp, const float* params);
int mhip_mlp_get_params(mhip_mlp* mlp, float* params);
int mhip_mlp_forward(mhip_mlp* mlp, int batch, const float* x, float* y);
int mhip_mlp_gradients(mhip_mlp* mlp, int batch, const float* x, const float* t, float* grad,
                       float* loss);
int mhip_mlp_train_step(mhip_mlp* mlp, int batch, const float* x, const float* t, float lr,
                        float* loss);

/* Learned starting density of a water cluster, built on the device: pair
 * descriptors, the network and the packed-density assembly all run in HBM and
 * the density stays there for the next mhip_scf with start_from_guess = 1.
 * The system must be uploaded and made of nw waters (O: 4 orbitals, H: 1).
 * waters: nw x 3 atom indices (O, H1, H2; 0-based); pairs (pair_a[m] >=
 * pair_b[m]): every atom with itself, the atoms of one molecule, and every
 * pair closer than 4 A (other blocks start at zero).  mlp: 96 -> ... -> 16.
 * x_out (npairs x 96, normalized descriptors) and y_out (npairs x 16, raw
 * network outputs) are filled when not NULL.  No counterpart in the
 * reference; the density plays the role of an OLDENS start
 * (src/SCF/iter.F90:151-159). */
typedef struct {
  const float* x_mean; /* 96 */
  const float* x_std;  /* 96 */
  const float* y_mean; /* 8 pair types x 16 */
  const float* y_std;  /* 8 x 16 */
} mhip_guess_model;
int mhip_guess_water(mhip_ctx* ctx, mhip_mlp* mlp, const mhip_guess_model* model, const int* nat,
                     const double* xyz, int nw, const int* waters, int npairs, const int* pair_a,
                     const int* pair_b, float* x_out, float* y_out);
/* Copy the device guess (mpack) to the host. */
int mhip_guess_density(mhip_ctx* ctx, double* p);

/* ---- diagnostics / tuning --------------------------------------------------
 * No counterpart in the reference.  kernel_variant: 1 = FP64 MFMA tiles
 * (default), 0 = vector-FMA tiles, <0 unchanged; max_inner: inner Jacobi
 * sweeps per 64x64 pair problem (<=0 unchanged); tol_iter: eigensolver
 * stopping tolerance inside the SCF loop (<=0 unchanged). */
int mhip_set_tuning(mhip_ctx* ctx, int kernel_variant, int max_inner, double tol_iter);
/* No counterpart in the reference.  Pair-solve schedule of the block-Jacobi
 * eigensolver for the whole process: 1 (default) = per step, 32 rounds over
 * the index pairs across the two blocks, with the pairs inside each block
 * rotated once per sweep; 0 = a full cyclic sweep over all 64 indices in
 * every step.  Other values leave it unchanged.  Returns the previous one. */
int mhip_set_jacobi_schedule(int schedule);
/* Block-Jacobi sweeps used by the last eigensolve of this context. */
int mhip_last_sweeps(mhip_ctx* ctx);
/* HIP-event timing of the block-Jacobi solve / A-update / C-update launches
 * on the context's stream.  Returns the accumulated milliseconds and launch
 * counts, the 64x64 tiles / block pairs the timed launches actually
 * processed and the 2x2 rotations of the timed pair solves; enable = 1 or 0
 * resets and switches timing on or off, -1 only reads. */
int mhip_kernel_timing(mhip_ctx* ctx, int enable, double* solve_ms, long long* n_solve,
                       double* apply_ms, long long* n_apply, double* applyc_ms,
                       long long* n_applyc, long long* active_tiles, long long* active_pairs,
                       long long* rotations);

/* Process-wide HIP-event timing of the other kernel classes (FP64 GEMM,
 * split-BF16 GEMM, packed density, fock2), on the stream each launch uses.
 * enable = 1 / 0 resets and switches it on / off; -1 fills, per class
 * (index < max_classes, at most 4), the summed milliseconds, launches,
 * algorithmic flops and algorithmic HBM bytes of the recorded launches. */
int mhip_kernel_stats(mhip_ctx* ctx, int enable, int max_classes, double* ms, long long* launches,
                      double* flops, double* bytes);
const char* mhip_kernel_class_name(int cls);

/* ---- PM7 Hamiltonian (host, once per geometry) ---------------------------
 * The producer of the loop's inputs (hcore, src/integrals/hcore.F90), for
 * the elements H, C, N, O.  Not on the per-iteration path. */
int mhip_pm7_sizes(int numat, const int* nat, int* norbs, long long* n2elec);
int mhip_pm7_hamiltonian(int numat, const int* nat, const double* coord, int charge,
                         int* nfirst, int* nlast, double* h, double* w, double* enuclr,
                         double* atheat, double* pdiag);
double mhip_ev_to_kcal(void);
/* Per-element PM7 values (parameters_C after calpar, src/models/calpar.F90):
 * out[0..8] = natorb, tore, am, ad, dd, aq, qq (atomic units), eisol (eV),
 * eheat (kcal/mol).  Elements 1, 6, 7, 8; -1 for any other. */
int mhip_pm7_element(int z, double* out);
/* Everything one atom pair contributes to hcore, for atom i (element zi at xi,
 * Angstrom) listed after atom j (zj at xj): w = the two-electron block
 * (limij x limkl, rotate's order), e1b/e2a = the electron-core attraction
 * blocks (packed, 10 each), enuc = the core-core term (eV), di = the
 * resonance block (4 x 4 row-major, orbitals of i down).  Replaces rotate and
 * h1elec for one pair (src/integrals/rotate.F90:16-124, h1elec.F90:16-84). */
int mhip_pm7_pair(int zi, int zj, const double* xi, const double* xj, double* w, double* e1b, double* e2a,
                  double* enuc, double* di);
/* ---- density restart files (MOPAC .den) -----------------------------------
 * The alpha density (packed, mpack) as den_in_out writes it with DENOUT /
 * DENOUTF and reads it with OLDENS (src/SCF/iter.F90:1133-1227, :151-159):
 * one Fortran sequential record (norbs, numat, pa) when formatted == 0,
 * list-directed text when formatted == 1.  mhip_den_read tries the form
 * asked for, then the other, and fails when the file's norbs or numat
 * differ ("Density file read in does not match current data set") or the
 * file is missing or short.  A restart starts mhip_scf through
 * mhip_scf_opts.p_init = 2 pa, as OLDENS sets p = 2 pa and pold = p. */
int mhip_den_write(const char* path, int norbs, int numat, const double* pa, int formatted);
int mhip_den_read(const char* path, int norbs, int numat, double* pa, int formatted);

/* PM7's post-SCF corrections of a discrete molecule (kcal/mol), the terms
 * post_scf_corrections adds to the SCF heat of formation for PM7
 * (src/corrections/post_scf_corrections.F90:112-114): *e_disp the damped
 * C6 dispersion (PM6_DH_Dispersion, H_bond_correction_PM6_DH_Dispersion.F90),
 * *e_hbond the DH+ hydrogen-bond term (Hydrogen_bond_corrections with
 * EH_plus), *n_hbonds the bonds below -1 kcal/mol (may be NULL).  Elements
 * 1, 6, 7, 8; coord is numat x 3 Angstrom. */
int mhip_pm7_corrections(int numat, const int* nat, const double* coord, double* e_disp, double* e_hbond,
                         int* n_hbonds);
/* The one-electron matrix on the MOZYME sparse storage (nijbo, mpack) and
 * the core-core energy as hcore_for_MOZYME forms them for a discrete system
 * (src/MOZYME/hcore_for_MOZYME.F90:100-260): NDDO terms for stored pairs,
 * with the resonance block zero beyond overlap_cutoff (CUTOFS, Angstrom,
 * default 7: fillij.F90:120-126, h1elec.F90:53) and from dipole_cutoff
 * (CUTOF1, diat.F90:97), outer2's point charge + dipole for the nijbo = -2
 * shell, outer1's point charge for the rest.  The near-pair integrals are the
 * conventional ones (mhip_pm7_hamiltonian's w without the other pairs'
 * blocks).  Returns -1 for an unsupported element, a nijbo beyond mpack or a
 * cutoff <= 0. */
int mhip_pm7_hamiltonian_mozyme(int numat, const int* nat, const double* coord, const int* nijbo,
                                long long mpack, double overlap_cutoff, double dipole_cutoff, double* h,
                                double* enuclr);
/* mhip_pm7_hamiltonian's atom-level outputs alone (nfirst, nlast, atheat,
 * pdiag), for systems whose pair terms are built sparsely. */
int mhip_pm7_atoms(int numat, const int* nat, int charge, int* nfirst, int* nlast, double* atheat,
                   double* pdiag);
/* The two-electron integrals of the pairs stored in nijbo (>= 0), in hcore
 * order (per atom i: its stored pairs j < i, then its one-centre block):
 * the rotate blocks hcore_for_MOZYME stores for near pairs
 * (hcore_for_MOZYME.F90:205-214); the far pairs' point-charge terms are
 * formed on the fly by the MOZYME Fock build, so none are stored here.
 * With w == NULL only *nw is set (the length); otherwise *nw is the capacity
 * on entry and the length on return. */
int mhip_pm7_integrals_mozyme(int numat, const int* nat, const double* coord, const int* nijbo,
                              long long* nw, double* w);

#ifdef __cplusplus
}
#endif
#endif /* MOPAC_HIP_H */
