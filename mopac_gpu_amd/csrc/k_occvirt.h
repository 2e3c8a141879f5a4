// Block pseudo-diagonalization: every occupied-virtual rotation of one
// pseudo-diagonalization pass (src/matrix/diag.F90:16-163) applied at once as
// GEMMs, followed by a Bjorck re-orthonormalization of each block.
#pragma once
#include "common.h"

namespace mhip {

struct OccVirtWork {
  DevBuf fd, wv, fmo, b, eo, ev, cw;
  void ensure(int n, int nocc, int ld);
  void release();
};

// One step on the n real rows/columns of C (ld x ld, column-major, columns
// 0..nocc-1 occupied, sorted as the last full diagonalization left them):
//   FMO = C_occ^T F C_virt,
//   B(v, j) = sin of the 2x2 rotation diag.F90 would apply to (j, v)
//             (angle from FMO(j, v) and eig(j) - eig(v)),
//   C_occ += C_virt B,  C_virt -= C_occ B^T   (simultaneous rotations),
//   one Bjorck step per block.
// eig is not changed (diag.F90 leaves it too).  Returns nothing on the host;
// everything stays on the stream.
void occvirt_step(hipStream_t s, OccVirtWork& w, int n, int ld, const double* fpacked, double* C,
                  int nocc, const double* eig);

// C(:, c0:c0+m) <- C (3I - C^T C) / 2 on the first n rows.
void bjorck_step(hipStream_t s, OccVirtWork& w, int n, int ld, double* C, int c0, int m);

}  // namespace mhip
