// Pseudo-diagonalization on the device (k_pseudo.hip).
#pragma once
#include "common.h"

namespace mhip {

struct PseudoWork {
  DevBuf fd, wv, fmo, flag, ab, cols, rot, counts, offs, part;
  void ensure(int n, int nocc, int ld);
  void release();
};

// C (n x n, leading dimension ld) is rotated in place; *nrot_out (host) gets
// the number of rotations applied when non-null.
void pseudo_diag_device(hipStream_t s, PseudoWork& w, int n, int ld, const double* fpacked,
                        double* C, int nocc, const double* eig, long long* nrot_out);

// Apply an explicit ordered list of (column a, column b, alpha, beta)
// rotations: a' = alpha a + beta b, b' = alpha b - beta a.
void apply_rotation_list(hipStream_t s, int n, int ld, double* C, const long long* nrot_dev,
                         const int2* cols, const double2* rot);

}  // namespace mhip
