// Split-BF16 GEMM on the BF16 matrix cores: C = alpha op(A) op(B) + beta C
// for FP64 operands, the precision form's contraction.
//
// Each FP64 operand element is split on its way into LDS as
// x = x0 + x1 + x2 (three BF16 values, x1 = bf16(x - x0), x2 = bf16(x - x0 - x1),
// 24 significant bits in all).  v_mfma_f32_16x16x16_bf16 forms the six
// products with i + j <= 2:
//   big  = x0 y0                                 (exact products, FP32 sum)
//   corr = x0 y1 + x1 y0 + x0 y2 + x1 y1 + x2 y0 (each ~2^-8 of big or less)
// into two FP32 accumulators, which are added into FP64 every KCHUNK values
// of k, so the FP32 rounding only acts on 32-term partial sums: the relative
// error is ~2^-24 per product and ~2^-24 sqrt(32) per chunk, not
// 2^-24 sqrt(K).  Reported error against FP64 is measured by the tests.
//
// v_mfma_f32_16x16x16_bf16 lane layout (64 lanes, l = lane id), measured on
// gfx950 with structured operands:
//   A (16x16): lane l holds A[l % 16][4 (l / 16) + v], v = 0..3
//   B (16x16): lane l holds B[4 (l / 16) + v][l % 16]
//   D (16x16): lane l holds D[4 (l / 16) + t][l % 16] in acc[t]
// so both operands are kept k-contiguous in LDS and every lane reads its four
// BF16 values with one 8-byte load.
#include "common.h"

namespace mhip {
namespace {

typedef short s4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int SB_TILE = 64;    // output tile edge
constexpr int SB_K = 32;       // k values per LDS stage (two MFMA k-steps)
constexpr int SB_NT = 256;     // 4 waves, 32x32 quadrant each
constexpr int KCHUNK = 32;     // k values per FP32 partial sum (one stage)
constexpr int LDK = SB_K + 4;  // padded k stride (shorts)

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ double bf16_to_d(unsigned short h) {
  return (double)__uint_as_float(((unsigned)h) << 16);
}

__device__ __forceinline__ void split3(double x, unsigned short& h0, unsigned short& h1,
                                       unsigned short& h2) {
  h0 = bf16_rne((float)x);
  const double r1 = x - bf16_to_d(h0);
  h1 = bf16_rne((float)r1);
  const double r2 = r1 - bf16_to_d(h1);
  h2 = bf16_rne((float)r2);
}

template <bool T>
__device__ __forceinline__ double ld_op(const double* p, int ld, int r, int k, int R, int K) {
  if (r >= R || k >= K) return 0.0;
  return T ? p[(size_t)k + (size_t)r * ld] : p[(size_t)r + (size_t)k * ld];
}

template <bool TA, bool TBT>
__global__ __launch_bounds__(SB_NT) void k_gemm_split_bf16(int M, int N, int K, double alpha,
                                                           const double* __restrict__ a, int lda,
                                                           const double* __restrict__ b, int ldb,
                                                           double beta, double* __restrict__ c,
                                                           int ldc) {
  // [plane][row or col][k]
  __shared__ __attribute__((aligned(16))) unsigned short As[3][SB_TILE][LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[3][SB_TILE][LDK];
  const int row0 = blockIdx.y * SB_TILE, col0 = blockIdx.x * SB_TILE;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int qr = 32 * (w >> 1), qc = 32 * (w & 1);
  const int li = lane & 15, lg = lane >> 4;
  f4_t big[4], corr[4];
  double tot[4][4];
  for (int i = 0; i < 4; ++i) {
    big[i] = f4_t{0.f, 0.f, 0.f, 0.f};
    corr[i] = f4_t{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 4; ++t) tot[i][t] = 0.0;
  }
  int since_flush = 0;
  for (int k0 = 0; k0 < K; k0 += SB_K) {
    // stage 64 x 32 of op(A) and 32 x 64 of op(B), split into three planes
#pragma unroll
    for (int r = 0; r < (SB_TILE * SB_K) / SB_NT; ++r) {
      const int e = tid + r * SB_NT;
      int mm, kk;
      if (TA) { kk = e % SB_K; mm = e / SB_K; } else { mm = e % SB_TILE; kk = e / SB_TILE; }
      unsigned short h0, h1, h2;
      split3(ld_op<TA>(a, lda, row0 + mm, k0 + kk, M, K), h0, h1, h2);
      As[0][mm][kk] = h0;
      As[1][mm][kk] = h1;
      As[2][mm][kk] = h2;
    }
#pragma unroll
    for (int r = 0; r < (SB_TILE * SB_K) / SB_NT; ++r) {
      const int e = tid + r * SB_NT;
      int nn, kk;
      // op(B)(k, n): B(k, n) at b[k + n ldb] (k fast) or B^T at b[n + k ldb] (n fast)
      if (TBT) { nn = e % SB_TILE; kk = e / SB_TILE; } else { kk = e % SB_K; nn = e / SB_K; }
      unsigned short h0, h1, h2;
      split3(ld_op<!TBT>(b, ldb, col0 + nn, k0 + kk, N, K), h0, h1, h2);
      Bs[0][nn][kk] = h0;
      Bs[1][nn][kk] = h1;
      Bs[2][nn][kk] = h2;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < SB_K; ks += 16) {
      s4_t av[2][3], bv[2][3];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          av[h][p] = *reinterpret_cast<const s4_t*>(&As[p][qr + 16 * h + li][ks + 4 * lg]);
          bv[h][p] = *reinterpret_cast<const s4_t*>(&Bs[p][qc + 16 * h + li][ks + 4 * lg]);
        }
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) {
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const int q = bi * 2 + bj;
          big[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[bi][0], bv[bj][0], big[q], 0, 0, 0);
          corr[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[bi][0], bv[bj][1], corr[q], 0, 0, 0);
          corr[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[bi][1], bv[bj][0], corr[q], 0, 0, 0);
          corr[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[bi][0], bv[bj][2], corr[q], 0, 0, 0);
          corr[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[bi][1], bv[bj][1], corr[q], 0, 0, 0);
          corr[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[bi][2], bv[bj][0], corr[q], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    since_flush += SB_K;
    if (since_flush >= KCHUNK || k0 + SB_K >= K) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          tot[q][t] += (double)big[q][t] + (double)corr[q][t];
        }
        big[q] = f4_t{0.f, 0.f, 0.f, 0.f};
        corr[q] = f4_t{0.f, 0.f, 0.f, 0.f};
      }
      since_flush = 0;
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gr = row0 + qr + 16 * bi + 4 * lg + t, gc = col0 + qc + 16 * bj + li;
        if (gr >= M || gc >= N) continue;
        double* dst = c + (size_t)gr + (size_t)gc * ldc;
        const double v = tot[bi * 2 + bj][t];
        *dst = (beta == 0.0) ? alpha * v : fma(beta, *dst, alpha * v);
      }
    }
  }
}

inline int cdiv_i(int a, int b) { return (a + b - 1) / b; }

}  // namespace

void gemm_split_bf16(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                     const double* a, int lda, const double* b, int ldb, double beta, double* c,
                     int ldc) {
  if (m <= 0 || n <= 0) return;
  const dim3 grid(cdiv_i(n, SB_TILE), cdiv_i(m, SB_TILE));
  const bool TA = (ta == 'T' || ta == 't'), TBT = (tb == 'T' || tb == 't');
  timed(s, K_GEMM_SPLIT, 2.0 * m * n * (double)k, 8.0 * ((double)m * k + (double)k * n + 2.0 * m * n), [&] {
  if (!TA && !TBT)
    hipLaunchKernelGGL((k_gemm_split_bf16<false, false>), grid, dim3(SB_NT), 0, s, m, n, k, alpha, a,
                       lda, b, ldb, beta, c, ldc);
  else if (TA && !TBT)
    hipLaunchKernelGGL((k_gemm_split_bf16<true, false>), grid, dim3(SB_NT), 0, s, m, n, k, alpha, a,
                       lda, b, ldb, beta, c, ldc);
  else if (!TA && TBT)
    hipLaunchKernelGGL((k_gemm_split_bf16<false, true>), grid, dim3(SB_NT), 0, s, m, n, k, alpha, a,
                       lda, b, ldb, beta, c, ldc);
  else
    hipLaunchKernelGGL((k_gemm_split_bf16<true, true>), grid, dim3(SB_NT), 0, s, m, n, k, alpha, a,
                       lda, b, ldb, beta, c, ldc);
  MHIP_LAUNCH_CHECK();
  });
}

}  // namespace mhip
