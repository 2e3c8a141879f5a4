// Block pseudo-diagonalization (GEMM form of src/matrix/diag.F90:16-163).
//
// diag.F90 applies the occupied-virtual 2x2 rotations one after another; all
// their angles come from one FMO = C_occ^T F C_virt and the eigenvalues of
// the last full diagonalization, so the whole pass is, to first order in the
// rotation angles, the single update
//     C_occ' = C_occ + C_virt B,   C_virt' = C_virt - C_occ B^T,
// with B(v, j) the sine of the (j, v) rotation.  Because C_occ^T C_virt = 0,
// the two blocks stay exactly orthogonal to each other; each block is then
// brought back to orthonormal columns by a Bjorck step C <- C (3I - C^T C)/2,
// whose error is second order in C^T C - I.  Every operation is a GEMM on the
// FP64 MFMA tiles; the fixed point (FMO = 0) is the same as diag.F90's.
#include "common.h"
#include "k_occvirt.h"

namespace mhip {
namespace {

constexpr int TT = 32;

// B(v, j) (ld nvirt) from FMO(j, v) (ld nocc): the rotation of diag.F90:139-153
__global__ __launch_bounds__(256) void k_ov_sines(int nocc, int nvirt, const double* __restrict__ fmo,
                                                  const double* __restrict__ eig,
                                                  double* __restrict__ b) {
  __shared__ double tile[TT][TT + 1];
  const int j0 = blockIdx.x * TT, v0 = blockIdx.y * TT;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < TT; r += 8) {
    const int j = j0 + tx, v = v0 + r;
    double beta = 0.0;
    if (j < nocc && v < nvirt) {
      const double c = fmo[(size_t)j + (size_t)v * nocc];
      const double d = eig[j] - eig[nocc + v];
      // diag.F90's beta = -sign(c) sqrt(1 - alpha^2), alpha^2 = (1 + |d|/s)/2,
      // s = sqrt(4c^2 + d^2), rewritten without the cancellation in
      // 1 - alpha^2 (which rounds every |beta| below ~1.5e-8 to 0 or 1.5e-8
      // and would stall the iteration there): beta = -c sqrt(2 / ((s + |d|) s)).
      const double sq = sqrt(4.0 * c * c + d * d);
      if (sq > 0.0) beta = -c * sqrt(2.0 / ((sq + fabs(d)) * sq));
    }
    tile[r][tx] = beta;  // [v][j]
  }
  __syncthreads();
  for (int r = ty; r < TT; r += 8) {
    const int j = j0 + r, v = v0 + tx;
    if (j < nocc && v < nvirt) b[(size_t)v + (size_t)j * nvirt] = tile[tx][r];
  }
}

__global__ void k_sub_identity(int m, double* __restrict__ e, int lde) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) e[(size_t)i + (size_t)i * lde] -= 1.0;
}

void copy_cols(hipStream_t s, int n, int ld, const double* src, double* dst, int ncols) {
  MHIP_CHECK(hipMemcpy2DAsync(dst, sizeof(double) * ld, src, sizeof(double) * ld, sizeof(double) * n,
                              ncols, hipMemcpyDeviceToDevice, s));
}

}  // namespace

void OccVirtWork::ensure(int n, int nocc, int ld) {
  const size_t nv = std::max(n - nocc, 1), no = std::max(nocc, 1);
  fd.ensure(sizeof(double) * (size_t)ld * ld);
  wv.ensure(sizeof(double) * (size_t)ld * nv);
  fmo.ensure(sizeof(double) * no * nv);
  b.ensure(sizeof(double) * no * nv);
  // m x m Gram blocks for any column block of C up to ld wide
  eo.ensure(sizeof(double) * (size_t)ld * ld);
  ev.ensure(sizeof(double) * (size_t)ld * ld);
  cw.ensure(sizeof(double) * (size_t)ld * ld);
}

void OccVirtWork::release() {
  for (DevBuf* x : {&fd, &wv, &fmo, &b, &eo, &ev, &cw}) x->release();
}

void bjorck_step(hipStream_t s, OccVirtWork& w, int n, int ld, double* C, int c0, int m) {
  if (m <= 0) return;
  MHIP_REQUIRE(c0 >= 0 && c0 + m <= ld && n <= ld, "bjorck_step: column block out of range");
  w.ensure(n, 0, ld);
  double* cb = C + (size_t)c0 * ld;
  double* cw = w.cw.as<double>() + (size_t)c0 * ld;
  // the occupied block (c0 == 0) and the virtual block use their own m x m buffer
  double* e = (c0 == 0) ? w.eo.as<double>() : w.ev.as<double>();
  copy_cols(s, n, ld, cb, cw, m);
  // E = C^T C - I
  gemm(s, 'T', 'N', m, m, n, 1.0, cw, ld, cw, ld, 0.0, e, m);
  hipLaunchKernelGGL(k_sub_identity, dim3((m + 255) / 256), dim3(256), 0, s, m, e, m);
  MHIP_LAUNCH_CHECK();
  // C <- C - C E / 2
  gemm(s, 'N', 'N', n, m, m, -0.5, cw, ld, e, m, 1.0, cb, ld);
}

void occvirt_step(hipStream_t s, OccVirtWork& w, int n, int ld, const double* fpacked, double* C,
                  int nocc, const double* eig) {
  const int nvirt = n - nocc;
  if (nocc <= 0 || nvirt <= 0) return;
  MHIP_REQUIRE(nocc <= ld && n <= ld, "occvirt_step: bad sizes");
  w.ensure(n, nocc, ld);
  double* fd = w.fd.as<double>();
  unpack_sym(s, n, fpacked, fd, ld, n, 0.0);
  double* cv = C + (size_t)nocc * ld;
  // Wv = F C_virt (n x nvirt), FMO = C_occ^T Wv (nocc x nvirt)
  gemm(s, 'N', 'N', n, nvirt, n, 1.0, fd, ld, cv, ld, 0.0, w.wv.as<double>(), ld);
  gemm(s, 'T', 'N', nocc, nvirt, n, 1.0, C, ld, w.wv.as<double>(), ld, 0.0, w.fmo.as<double>(), nocc);
  hipLaunchKernelGGL(k_ov_sines, dim3((nocc + TT - 1) / TT, (nvirt + TT - 1) / TT), dim3(256), 0, s,
                     nocc, nvirt, w.fmo.as<double>(), eig, w.b.as<double>());
  MHIP_LAUNCH_CHECK();
  // simultaneous rotations, both from the old C
  double* cw = w.cw.as<double>();
  copy_cols(s, n, ld, C, cw, n);
  gemm(s, 'N', 'N', n, nocc, nvirt, 1.0, cw + (size_t)nocc * ld, ld, w.b.as<double>(), nvirt, 1.0, C,
       ld);
  gemm(s, 'N', 'T', n, nvirt, nocc, -1.0, cw, ld, w.b.as<double>(), nvirt, 1.0, cv, ld);
  bjorck_step(s, w, n, ld, C, 0, nocc);
  bjorck_step(s, w, n, ld, C, nocc, nvirt);
}

}  // namespace mhip
