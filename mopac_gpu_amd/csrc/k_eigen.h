// Block-Jacobi symmetric eigensolver (k_eigen.hip).
#pragma once
#include "common.h"

namespace mhip {

struct JacobiPlan {
  int n = 0, npad = 0, nb = 0, npairs = 0;
  int nsteps = 0;        // steps per sweep
  bool grouped = false;  // pairs only within [0, split) and [split, n)
  int split = 0;
  int npo = 0;  // pairs of the first group in every step (npairs when not grouped)
  int max_inner = 2;  // inner Jacobi sweeps per 64x64 pair problem
  // Optional HIP-event timing of the solve and A-update launches.
  bool timing = false;
  double t_solve_ms = 0.0, t_apply_ms = 0.0, t_applyc_ms = 0.0;
  long long n_solve = 0, n_apply = 0, n_applyc = 0;
  long long active_tiles = 0, active_pairs = 0;  // summed over timed launches
  long long active_rotations = 0;                 // 2x2 rotations of the timed solves
  std::vector<int> intra_host;                    // host copy of intra
  std::vector<hipEvent_t> ev;
  DevBuf pairs;  // nsteps x npairs x 2 block indices
  DevBuf qbuf;   // npairs x 64 x 64 rotations of the current step
  DevBuf part;   // reduction scratch
  DevBuf perm;   // sort permutation
  DevBuf active; // per pair: 1 when Q != I this step
  DevBuf intra;  // nsteps x npairs: 1 where the cross-round solve also
                 // rotates inside the two blocks (once per sweep of a group)
  DevBuf counts; // active pairs per step (timing mode)
  // C <- C Q runs on its own stream, behind by up to a sweep: Q and the
  // active flags have a slot per step, and events order each slot's C update
  // after its step's A update and before the slot's next solve
  hipStream_t cstream = nullptr;
  std::vector<hipEvent_t> ev_solved, ev_cdone;
  std::vector<char> cdone_pending;
  hipEvent_t ev_join = nullptr;
  // split > 0 (a multiple of 32 leaving an even number of blocks on both
  // sides): rotations only within the leading split columns and within the
  // rest, for a matrix already block-diagonal in that partition.
  void build(int n, int split = 0);
  void release() {
    if (cstream) (void)hipStreamSynchronize(cstream);
    pairs.release(); qbuf.release(); part.release(); perm.release(); active.release();
    intra.release(); counts.release();
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    for (auto e : ev_solved) (void)hipEventDestroy(e);
    for (auto e : ev_cdone) (void)hipEventDestroy(e);
    ev_solved.clear();
    ev_cdone.clear();
    cdone_pending.clear();
    if (ev_join) (void)hipEventDestroy(ev_join);
    ev_join = nullptr;
    if (cstream) (void)hipStreamDestroy(cstream);
    cstream = nullptr;
  }
};

// Pair-solve schedule of every plan in the process: 1 (default) = cross
// rounds with one intra-block pass per sweep, 0 = a full 64-index cyclic
// sweep in every step.
extern int g_jacobi_schedule;

void jacobi_set_identity(hipStream_t s, int npad, double* C, int ldc);
// Sweeps until max|offdiag| <= tol * max(1, max|diag|).  Returns 0 when
// converged, 1 when max_sweeps ran out.
int jacobi_sweeps(hipStream_t s, JacobiPlan& plan, double* A, double* C, double tol, int max_sweeps,
                  double* host_scal, double* dev_scal, int* sweeps_done);
// eig_dev[k] = k-th smallest diagonal entry, Cout[:,k] = C[:,perm[k]].
void jacobi_sort(hipStream_t s, JacobiPlan& plan, const double* A, const double* C, double* Cout,
                 double* eig_dev);
void phase_lock(hipStream_t s, int n, double* C, int ldc);

}  // namespace mhip
