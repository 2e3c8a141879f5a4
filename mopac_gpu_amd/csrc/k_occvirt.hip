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

// colmax[blockIdx.x] = max over this block's occupied columns j of sum_v B(v,j)^2
__global__ __launch_bounds__(256) void k_b_colnorm(int nocc, int nvirt, const double* __restrict__ b,
                                                   double* __restrict__ part) {
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double mx = 0.0;
  for (int j = blockIdx.x * 4 + w; j < nocc; j += gridDim.x * 4) {
    double s = 0.0;
    for (int v = lane; v < nvirt; v += 64) {
      const double x = b[(size_t)v + (size_t)j * nvirt];
      s += x * x;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    mx = fmax(mx, s);
  }
  if (lane == 0) red[w] = mx;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// B *= min(1, cap / max column norm): bounds the step so that the Bjorck
// steps stay in their region of convergence far from self-consistency
__global__ __launch_bounds__(256) void k_b_scale(long long len, int nparts, const double* __restrict__ part,
                                                 double cap, double* __restrict__ b) {
  double mx = 0.0;
  for (int i = 0; i < nparts; ++i) mx = fmax(mx, part[i]);
  const double nrm = sqrt(mx);
  if (!(nrm > cap)) return;
  const double sc = cap / nrm;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < len; i += (long long)gridDim.x * 256) b[i] *= sc;
}

// eig(nocc + v) = C_v(:, v) . (F C_v)(:, v): the virtual levels of the current
// F (level shift and SCF changes included) in the current virtual orbitals
__global__ __launch_bounds__(256) void k_virt_rayleigh(int n, int nocc, int nvirt, int ld,
                                                       const double* __restrict__ cv,
                                                       const double* __restrict__ wv,
                                                       double* __restrict__ eig) {
  const int lane = threadIdx.x & 63;
  const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (v >= nvirt) return;
  double s = 0.0;
  for (int i = lane; i < n; i += 64) s += cv[(size_t)i + (size_t)v * ld] * wv[(size_t)i + (size_t)v * ld];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) eig[nocc + v] = s;
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
  b2.ensure(sizeof(double) * no * nv);
  // m x m Gram blocks for any column block of C up to ld wide
  eo.ensure(sizeof(double) * (size_t)ld * ld);
  ev.ensure(sizeof(double) * (size_t)ld * ld);
  cw.ensure(sizeof(double) * (size_t)ld * ld);
  part.ensure(sizeof(double) * 256);
}

void OccVirtWork::release() {
  for (DevBuf* x : {&fd, &wv, &fmo, &b, &b2, &eo, &ev, &cw, &part}) x->release();
}

void bjorck_step(hipStream_t s, OccVirtWork& w, int n, int ld, double* C, int c0, int m, bool split) {
  if (m <= 0) return;
  MHIP_REQUIRE(c0 >= 0 && c0 + m <= ld && n <= ld, "bjorck_step: column block out of range");
  w.ensure(n, 0, ld);
  double* cb = C + (size_t)c0 * ld;
  double* cw = w.cw.as<double>() + (size_t)c0 * ld;
  // the occupied block (c0 == 0) and the virtual block use their own m x m buffer
  double* e = (c0 == 0) ? w.eo.as<double>() : w.ev.as<double>();
  copy_cols(s, n, ld, cb, cw, m);
  // E = C^T C - I (symmetric: lower-triangle tiles, then mirrored)
  gemm_sym(s, 'T', 'N', m, n, 1.0, cw, ld, cw, ld, 0.0, e, m);
  hipLaunchKernelGGL(k_sub_identity, dim3((m + 255) / 256), dim3(256), 0, s, m, e, m);
  MHIP_LAUNCH_CHECK();
  // C <- C - C E / 2 (E is small: the split product's error is second order)
  if (split) gemm_split_bf16(s, 'N', 'N', n, m, m, -0.5, cw, ld, e, m, 1.0, cb, ld);
  else gemm(s, 'N', 'N', n, m, m, -0.5, cw, ld, e, m, 1.0, cb, ld);
}

void occvirt_step(hipStream_t s, OccVirtWork& w, int n, int ld, const double* fpacked, double* C,
                  int nocc, double* eig, bool split, int bjorck_steps, bool project) {
  const int nvirt = n - nocc;
  if (nocc <= 0 || nvirt <= 0) return;
  MHIP_REQUIRE(nocc <= ld && n <= ld, "occvirt_step: bad sizes");
  w.ensure(n, nocc, ld);
  double* fd = w.fd.as<double>();
  unpack_sym(s, n, fpacked, fd, ld, n, 0.0);
  double* cv = C + (size_t)nocc * ld;
  // Wv = F C_virt (n x nvirt), FMO = C_occ^T Wv (nocc x nvirt)
  auto mm = [&](char ta, char tb, int m, int nn, int k, double al, const double* a, int lda,
                const double* b, int ldb, double be, double* c, int ldc) {
    if (split) gemm_split_bf16(s, ta, tb, m, nn, k, al, a, lda, b, ldb, be, c, ldc);
    else gemm(s, ta, tb, m, nn, k, al, a, lda, b, ldb, be, c, ldc);
  };
  mm('N', 'N', n, nvirt, n, 1.0, fd, ld, cv, ld, 0.0, w.wv.as<double>(), ld);
  mm('T', 'N', nocc, nvirt, n, 1.0, C, ld, w.wv.as<double>(), ld, 0.0, w.fmo.as<double>(), nocc);
  hipLaunchKernelGGL(k_virt_rayleigh, dim3((nvirt + 3) / 4), dim3(256), 0, s, n, nocc, nvirt, ld, cv,
                     w.wv.as<double>(), eig);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ov_sines, dim3((nocc + TT - 1) / TT, (nvirt + TT - 1) / TT), dim3(256), 0, s,
                     nocc, nvirt, w.fmo.as<double>(), eig, w.b.as<double>());
  MHIP_LAUNCH_CHECK();
  constexpr int kParts = 256;
  hipLaunchKernelGGL(k_b_colnorm, dim3(kParts), dim3(256), 0, s, nocc, nvirt, w.b.as<double>(),
                     w.part.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_b_scale, dim3(1024), dim3(256), 0, s, (long long)nocc * nvirt, kParts,
                     w.part.as<double>(), kMaxStep, w.b.as<double>());
  MHIP_LAUNCH_CHECK();
  // simultaneous rotations, both from the old C
  double* cw = w.cw.as<double>();
  copy_cols(s, n, ld, C, cw, n);
  const double* bv = w.b.as<double>();
  if (project) {
    // b2 = B + C_virt^T C_occ (nvirt x nocc, FP64): the virtual update then
    // also takes out the virtual columns' overlap with the occupied ones
    MHIP_CHECK(hipMemcpyAsync(w.b2.ptr, w.b.ptr, sizeof(double) * (size_t)nocc * nvirt,
                              hipMemcpyDeviceToDevice, s));
    gemm(s, 'T', 'N', nvirt, nocc, n, 1.0, cw + (size_t)nocc * ld, ld, cw, ld, 1.0, w.b2.as<double>(), nvirt);
    bv = w.b2.as<double>();
  }
  mm('N', 'N', n, nocc, nvirt, 1.0, cw + (size_t)nocc * ld, ld, w.b.as<double>(), nvirt, 1.0, C, ld);
  mm('N', 'T', n, nvirt, nocc, -1.0, cw, ld, bv, nvirt, 1.0, cv, ld);
  for (int it = 0; it < bjorck_steps; ++it) {
    bjorck_step(s, w, n, ld, C, 0, nocc, split);
    bjorck_step(s, w, n, ld, C, nocc, nvirt, split);
  }
}

void orthogonalize_virtual(hipStream_t s, OccVirtWork& w, int n, int ld, double* C, int nocc, int bjorck_steps) {
  const int nvirt = n - nocc;
  if (nocc <= 0 || nvirt <= 0) return;
  w.ensure(n, nocc, ld);
  double* cv = C + (size_t)nocc * ld;
  // G = C_occ^T C_virt (nocc x nvirt), then C_virt -= C_occ G
  gemm(s, 'T', 'N', nocc, nvirt, n, 1.0, C, ld, cv, ld, 0.0, w.b2.as<double>(), nocc);
  gemm(s, 'N', 'N', n, nvirt, nocc, -1.0, C, ld, w.b2.as<double>(), nocc, 1.0, cv, ld);
  for (int it = 0; it < bjorck_steps; ++it) bjorck_step(s, w, n, ld, C, nocc, nvirt);
}

}  // namespace mhip

namespace mhip {
namespace {

__global__ __launch_bounds__(64) void k_fragment_eigen(const double* __restrict__ fd, int ld,
                                                       const int* __restrict__ fptr,
                                                       const int* __restrict__ forb,
                                                       const int* __restrict__ focc,
                                                       const int* __restrict__ occ_off,
                                                       const int* __restrict__ virt_off,
                                                       double* __restrict__ C,
                                                       double* __restrict__ eig) {
  constexpr int M = kMaxFragOrbs;
  __shared__ double a[M][M + 1], v[M][M + 1];
  __shared__ double cs[2];
  __shared__ int order[M];
  __shared__ int done;
  const int f = blockIdx.x, tid = threadIdx.x;
  const int o0 = fptr[f], nf = fptr[f + 1] - o0;
  for (int e = tid; e < M * M; e += 64) {
    const int r = e / M, c = e % M;
    a[r][c] = (r < nf && c < nf) ? fd[(size_t)forb[o0 + r] + (size_t)forb[o0 + c] * ld] : 0.0;
    v[r][c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 50; ++sweep) {
    if (tid == 0) {
      double off = 0.0, dg = 0.0;
      for (int r = 0; r < nf; ++r) {
        dg = fmax(dg, fabs(a[r][r]));
        for (int c = 0; c < r; ++c) off = fmax(off, fabs(a[r][c]));
      }
      done = off <= 1.0e-15 * fmax(dg, 1.0);
    }
    __syncthreads();
    if (done) break;
    for (int p = 0; p < nf - 1; ++p) {
      for (int q = p + 1; q < nf; ++q) {
        if (tid == 0) {
          const double app = a[p][p], aqq = a[q][q], apq = a[p][q];
          double c = 1.0, sn = 0.0;
          if (fabs(apq) > 1.0e-300) {
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            sn = t * c;
          }
          cs[0] = c;
          cs[1] = sn;
        }
        __syncthreads();
        const double c = cs[0], sn = cs[1];
        if (tid < nf) {  // rows p, q
          const double x = a[p][tid], y = a[q][tid];
          a[p][tid] = c * x - sn * y;
          a[q][tid] = sn * x + c * y;
        }
        __syncthreads();
        if (tid < nf) {  // columns p, q of a and v
          const double x = a[tid][p], y = a[tid][q];
          a[tid][p] = c * x - sn * y;
          a[tid][q] = sn * x + c * y;
          const double u = v[tid][p], w = v[tid][q];
          v[tid][p] = c * u - sn * w;
          v[tid][q] = sn * u + c * w;
        }
        __syncthreads();
      }
    }
  }
  if (tid == 0) {  // ascending order of the eigenvalues
    for (int k = 0; k < nf; ++k) order[k] = k;
    for (int k = 1; k < nf; ++k) {
      const int key = order[k];
      int m = k - 1;
      while (m >= 0 && a[order[m]][order[m]] > a[key][key]) {
        order[m + 1] = order[m];
        --m;
      }
      order[m + 1] = key;
    }
  }
  __syncthreads();
  const int no = focc[f];
  for (int e = tid; e < nf * nf; e += 64) {
    const int k = e / nf, r = e % nf;  // k-th eigenvector, row r
    const int col = (k < no) ? occ_off[f] + k : virt_off[f] + (k - no);
    C[(size_t)forb[o0 + r] + (size_t)col * ld] = v[r][order[k]];
  }
  if (tid < nf) {
    const int k = tid;
    const int col = (k < no) ? occ_off[f] + k : virt_off[f] + (k - no);
    eig[col] = a[order[k]][order[k]];
  }
}

}  // namespace

void fragment_eigen(hipStream_t s, int nfrag, const double* fd, int ld, const int* fptr,
                    const int* forb, const int* focc, const int* occ_off, const int* virt_off,
                    double* C, double* eig) {
  if (nfrag <= 0) return;
  hipLaunchKernelGGL(k_fragment_eigen, dim3(nfrag), dim3(64), 0, s, fd, ld, fptr, forb, focc, occ_off,
                     virt_off, C, eig);
  MHIP_LAUNCH_CHECK();
}

}  // namespace mhip
