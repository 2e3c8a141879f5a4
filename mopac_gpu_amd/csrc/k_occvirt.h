// Block pseudo-diagonalization: every occupied-virtual rotation of one
// pseudo-diagonalization pass (src/matrix/diag.F90:16-163) applied at once as
// GEMMs, followed by a Bjorck re-orthonormalization of each block.
#pragma once
#include "common.h"

namespace mhip {

struct OccVirtWork {
  DevBuf fd, wv, fmo, b, b2, eo, ev, cw, part;
  void ensure(int n, int nocc, int ld);
  void release();
};

// One step on the n real rows/columns of C (ld x ld, column-major, columns
// 0..nocc-1 occupied, sorted as the last full diagonalization left them):
//   FMO = C_occ^T F C_virt,
//   B(v, j) = sin of the 2x2 rotation diag.F90 would apply to (j, v)
//             (angle from FMO(j, v) and eig(j) - eig(v)),
//   C_occ += C_virt B,  C_virt -= C_occ B^T   (simultaneous rotations),
//   B scaled down if any column norm exceeds kMaxStep,
//   bjorck_steps Bjorck steps per block.
// The virtual levels eig(nocc..n-1) are refreshed first as Rayleigh quotients
// C_v^T F C_v of the current F, so the angles see the level shift iter applies
// to the virtual space (diag.F90 keeps the levels of the last diagonalization;
// here the occupied ones are kept).  Everything stays on the stream.  split = true runs the contractions on the
// split-BF16 GEMM (the precision form); C^T C - I is always formed in FP64.
// project = true also removes the virtual columns' overlap with the
// occupied ones in the same update: C_virt -= C_occ (B^T + C_occ^T C_virt),
// the overlap formed in FP64 from the old C.  The simultaneous rotations keep
// C_occ^T C_virt = 0 only to first order in the blocks' own orthonormality
// error, so without it that overlap accumulates over the iterations and the
// fixed point FMO = 0 no longer pins the occupied space to FP64 precision.
constexpr double kMaxStep = 0.3;
void occvirt_step(hipStream_t s, OccVirtWork& w, int n, int ld, const double* fpacked, double* C,
                  int nocc, double* eig, bool split, int bjorck_steps, bool project = false);

// The virtual columns made orthogonal to the occupied ones and then
// orthonormal among themselves: C_virt -= C_occ (C_occ^T C_virt), then
// `bjorck_steps` Bjorck steps on C_virt.  The occupied columns (and so the
// density) are left as they are.
void orthogonalize_virtual(hipStream_t s, OccVirtWork& w, int n, int ld, double* C, int nocc, int bjorck_steps);

// C(:, c0:c0+m) <- C (3I - C^T C) / 2 on the first n rows.
void bjorck_step(hipStream_t s, OccVirtWork& w, int n, int ld, double* C, int c0, int m,
                 bool split = false);

}  // namespace mhip

namespace mhip {

// Starting orbitals from fragments (molecules of a cluster): each fragment's
// diagonal block of the dense F (ld) is diagonalized on its own (one
// workgroup per fragment, cyclic Jacobi in LDS, nf <= kMaxFragOrbs), and the
// block-diagonal eigenvectors become C: fragment f's lowest focc[f]
// eigenvectors go to columns occ_off[f].., the rest to virt_off[f]..; eig
// follows the same column order.  C must be zero on entry outside the
// padding identity.
constexpr int kMaxFragOrbs = 16;
void fragment_eigen(hipStream_t s, int nfrag, const double* fd, int ld, const int* fptr,
                    const int* forb, const int* focc, const int* occ_off, const int* virt_off,
                    double* C, double* eig);

}  // namespace mhip
