// Pseudo-diagonalization (Stewart, Csaszar & Pulay 1982) as in
// src/matrix/diag.F90:16-163:
//   FMO = C_occ^T F C_virt (two GEMMs),
//   tiny = 0.05 max|FMO|,
//   for every (virtual i, occupied j) in that order whose |FMO| >= tiny and
//   |FMO/(e_j - e_i)| >= 1.5e-7, rotate columns j and i of C.
// FMO is not updated during the rotations, so every rotation angle is known
// up front: one kernel computes all of them, a scan compacts the accepted
// ones in order, and the rotation kernel keeps a slab of rows of C in LDS and
// applies the whole ordered list to it (rows are independent, so this is
// exactly the sequential result).
#include "common.h"
#include "k_pseudo.h"

namespace mhip {
namespace {

constexpr int SCAN_BLOCK = 1024;

__global__ __launch_bounds__(256) void k_absmax_partial(long long len, const double* __restrict__ x,
                                                        double* __restrict__ part) {
  double m = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
    m = fmax(m, fabs(x[i]));
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_absmax_final(int nb, const double* __restrict__ part,
                                                      double* __restrict__ out) {
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) m = fmax(m, part[i]);
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = 0.05 * red[0];
}

// Rotation parameters for every (virtual, occupied) pair, linear index
// ij = v * nocc + j, exactly as diag.F90:130-153 computes them.
__global__ void k_rot_params(int nocc, int nvirt, const double* __restrict__ fmo,
                             const double* __restrict__ eig, const double* __restrict__ tiny_dev,
                             double bigeps, int* __restrict__ flag, double2* __restrict__ ab) {
  const long long len = (long long)nocc * nvirt;
  const double tiny = tiny_dev[0];
  for (long long ij = blockIdx.x * (long long)blockDim.x + threadIdx.x; ij < len;
       ij += (long long)gridDim.x * blockDim.x) {
    const int v = (int)(ij / nocc), j = (int)(ij % nocc);
    const double c = fmo[ij];
    int ok = 0;
    double alpha = 1.0, beta = 0.0;
    if (!(fabs(c) < tiny)) {
      const double a = eig[j], b = eig[nocc + v];
      const double d = a - b;
      if (!(fabs(c / d) < bigeps)) {
        const double e = (d >= 0.0 ? 1.0 : -1.0) * sqrt(4.0 * c * c + d * d);
        alpha = sqrt(0.5 * (1.0 + d / e));
        beta = -(c >= 0.0 ? 1.0 : -1.0) * sqrt(1.0 - alpha * alpha);
        ok = 1;
      }
    }
    flag[ij] = ok;
    ab[ij] = make_double2(alpha, beta);
  }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_counts(long long len, const int* __restrict__ flag,
                                                            int* __restrict__ counts) {
  const long long i = blockIdx.x * (long long)SCAN_BLOCK + threadIdx.x;
  __shared__ int red[SCAN_BLOCK];
  red[threadIdx.x] = (i < len) ? flag[i] : 0;
  __syncthreads();
  for (int w = SCAN_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// exclusive scan of the per-block counts (single workgroup), total in offs[nb]
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_offsets(int nb, const int* __restrict__ counts,
                                                             long long* __restrict__ offs) {
  __shared__ long long buf[SCAN_BLOCK];
  long long carry = 0;
  for (int base = 0; base < nb; base += SCAN_BLOCK) {
    const int i = base + threadIdx.x;
    const long long v = (i < nb) ? counts[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
      const long long t = (threadIdx.x >= off) ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) offs[i] = carry + buf[threadIdx.x] - v;
    const long long tot = buf[SCAN_BLOCK - 1];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) offs[nb] = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_scatter(long long len, int nocc,
                                                             const int* __restrict__ flag,
                                                             const double2* __restrict__ ab,
                                                             const long long* __restrict__ offs,
                                                             int2* __restrict__ cols,
                                                             double2* __restrict__ rot) {
  const long long i = blockIdx.x * (long long)SCAN_BLOCK + threadIdx.x;
  __shared__ int buf[SCAN_BLOCK];
  const int v = (i < len) ? flag[i] : 0;
  buf[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    const int t = (threadIdx.x >= off) ? buf[threadIdx.x - off] : 0;
    __syncthreads();
    buf[threadIdx.x] += t;
    __syncthreads();
  }
  if (v) {
    const long long dst = offs[blockIdx.x] + buf[threadIdx.x] - 1;
    const int vv = (int)(i / nocc), j = (int)(i % nocc);
    cols[dst] = make_int2(j, nocc + vv);
    rot[dst] = ab[i];
  }
}

// Apply the ordered rotation list to `rows` consecutive rows of C held in LDS.
__global__ __launch_bounds__(64) void k_apply_rotations(int n, int ldc, double* __restrict__ C,
                                                        int rows, const long long* __restrict__ total,
                                                        const int2* __restrict__ cols,
                                                        const double2* __restrict__ rot) {
  extern __shared__ double slab[];  // rows x n, row-major
  const int row0 = blockIdx.x * rows;
  const int nr = min(rows, n - row0);
  for (int e = threadIdx.x; e < nr * n; e += 64) {
    const int r = e % nr, c = e / nr;
    slab[(size_t)r * n + c] = C[(size_t)(row0 + r) + (size_t)c * ldc];
  }
  __syncthreads();
  const long long nrot = *total;
  if (threadIdx.x < nr) {
    double* row = slab + (size_t)threadIdx.x * n;
    for (long long k = 0; k < nrot; ++k) {
      const int2 jc = cols[k];
      const double2 r = rot[k];
      const double a = row[jc.x], b = row[jc.y];
      row[jc.x] = r.x * a + r.y * b;
      row[jc.y] = r.x * b - r.y * a;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nr * n; e += 64) {
    const int r = e % nr, c = e / nr;
    C[(size_t)(row0 + r) + (size_t)c * ldc] = slab[(size_t)r * n + c];
  }
}

}  // namespace

void PseudoWork::ensure(int n, int nocc, int ld) {
  const long long nvirt = n - nocc;
  const long long len = (long long)nocc * nvirt;
  fd.ensure(sizeof(double) * (size_t)ld * ld);
  wv.ensure(sizeof(double) * (size_t)ld * std::max<long long>(nvirt, 1));
  fmo.ensure(sizeof(double) * std::max<long long>(len, 1));
  flag.ensure(sizeof(int) * std::max<long long>(len, 1));
  ab.ensure(sizeof(double2) * std::max<long long>(len, 1));
  cols.ensure(sizeof(int2) * std::max<long long>(len, 1));
  rot.ensure(sizeof(double2) * std::max<long long>(len, 1));
  const long long nblk = (len + SCAN_BLOCK - 1) / SCAN_BLOCK + 1;
  counts.ensure(sizeof(int) * nblk);
  offs.ensure(sizeof(long long) * (nblk + 1));
  part.ensure(sizeof(double) * 1024);
}

void PseudoWork::release() {
  for (DevBuf* b : {&fd, &wv, &fmo, &flag, &ab, &cols, &rot, &counts, &offs, &part}) b->release();
}

void pseudo_diag_device(hipStream_t s, PseudoWork& w, int n, int ld, const double* fpacked,
                        double* C, int nocc, const double* eig, long long* nrot_out) {
  const int nvirt = n - nocc;
  if (nvirt <= 0 || nocc <= 0) {
    if (nrot_out) *nrot_out = 0;
    return;
  }
  w.ensure(n, nocc, ld);
  unpack_sym(s, n, fpacked, w.fd.as<double>(), ld, n, 0.0);
  // Wv = F C_virt (n x nvirt), FMO = C_occ^T Wv (nocc x nvirt)
  gemm(s, 'N', 'N', n, nvirt, n, 1.0, w.fd.as<double>(), ld, C + (size_t)nocc * ld, ld, 0.0,
       w.wv.as<double>(), ld);
  gemm(s, 'T', 'N', nocc, nvirt, n, 1.0, C, ld, w.wv.as<double>(), ld, 0.0, w.fmo.as<double>(), nocc);
  const long long len = (long long)nocc * nvirt;
  const int nb = (int)std::min<long long>(1024, (len + 255) / 256);
  hipLaunchKernelGGL(k_absmax_partial, dim3(nb), dim3(256), 0, s, len, w.fmo.as<double>(),
                     w.part.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_absmax_final, dim3(1), dim3(256), 0, s, nb, w.part.as<double>(),
                     w.part.as<double>() + 1023);
  MHIP_LAUNCH_CHECK();
  const int pb = (int)std::min<long long>(8192, (len + 255) / 256);
  hipLaunchKernelGGL(k_rot_params, dim3(pb), dim3(256), 0, s, nocc, nvirt, w.fmo.as<double>(), eig,
                     w.part.as<double>() + 1023, 1.5e-7, w.flag.as<int>(), w.ab.as<double2>());
  MHIP_LAUNCH_CHECK();
  const int nblk = (int)((len + SCAN_BLOCK - 1) / SCAN_BLOCK);
  hipLaunchKernelGGL(k_scan_counts, dim3(nblk), dim3(SCAN_BLOCK), 0, s, len, w.flag.as<int>(),
                     w.counts.as<int>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(SCAN_BLOCK), 0, s, nblk, w.counts.as<int>(),
                     w.offs.as<long long>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_scatter, dim3(nblk), dim3(SCAN_BLOCK), 0, s, len, nocc,
                     w.flag.as<int>(), w.ab.as<double2>(), w.offs.as<long long>(),
                     w.cols.as<int2>(), w.rot.as<double2>());
  MHIP_LAUNCH_CHECK();
  const int rows = std::max(1, std::min(64, 19456 / n));
  const size_t lds = sizeof(double) * (size_t)rows * n;
  static bool attr_set = false;
  if (!attr_set) {
    MHIP_CHECK(hipFuncSetAttribute((const void*)k_apply_rotations,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_apply_rotations, dim3((n + rows - 1) / rows), dim3(64), lds, s, n, ld, C,
                     rows, w.offs.as<long long>() + nblk, w.cols.as<int2>(), w.rot.as<double2>());
  MHIP_LAUNCH_CHECK();
  if (nrot_out) {
    MHIP_CHECK(hipMemcpyAsync(nrot_out, w.offs.as<long long>() + nblk, sizeof(long long),
                              hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipStreamSynchronize(s));
  }
}

void apply_rotation_list(hipStream_t s, int n, int ld, double* C, const long long* nrot_dev,
                         const int2* cols, const double2* rot) {
  const int rows = std::max(1, std::min(64, 19456 / n));
  const size_t lds = sizeof(double) * (size_t)rows * n;
  MHIP_CHECK(hipFuncSetAttribute((const void*)k_apply_rotations,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  hipLaunchKernelGGL(k_apply_rotations, dim3((n + rows - 1) / rows), dim3(64), lds, s, n, ld, C,
                     rows, nrot_dev, cols, rot);
  MHIP_LAUNCH_CHECK();
}

}  // namespace mhip
