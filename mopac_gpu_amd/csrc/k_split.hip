// Split-BF16 GEMM on the BF16 matrix cores: C = alpha op(A) op(B) + beta C
// for FP64 operands, the precision form's contraction.
//
// Each FP64 operand element is split once, before the product, as
// x = x0 + x1 + x2 (three BF16 values, x1 = bf16(x - x0), x2 = bf16(x - x0 - x1),
// 24 significant bits in all).  The BF16 MFMA forms the six products with
// i + j <= 2:
//   big  = x0 y0                                 (exact products, FP32 sum)
//   corr = x0 y1 + x1 y0 + x0 y2 + x1 y1 + x2 y0 (each ~2^-8 of big or less)
// into two FP32 accumulators, which are added into FP64 every 64 values of
// k, so the FP32 rounding only acts on 64-term partial sums: the relative
// error is ~2^-24 per product and ~2^-24 sqrt(64) per chunk, not
// 2^-24 sqrt(K).  Reported error against FP64 is measured by the tests.
//
// MFMA lane layout (64 lanes, l = lane id; measured on gfx950 for the 16-deep
// form, and the 32-deep form keeps rows, columns and D alike):
//   A (16 x kdepth): lane l holds row l % 16, a contiguous group of k
//   B (kdepth x 16): lane l holds column l % 16, the same k group
//   D (16 x 16):     lane l holds D[4 (l / 16) + t][l % 16] in acc[t]
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.h"

namespace mhip {
namespace {

typedef float f4_t __attribute__((ext_vector_type(4)));


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

inline int cdiv_i(int a, int b) { return (a + b - 1) / b; }

// ---- planes form: operands split once, then a large-tile product ----------
//
// k_split_planes writes op(X) (R x K) as three BF16 planes [plane][r][k] with
// k contiguous (row stride kp), zero-padded to rp rows and kp columns, so
// the product kernel needs no bounds checks and reads 16 bytes per lane.
// k_gemm_split_planes: 128 x 64 output tile per workgroup (eight waves of
// 32 x 32), k in steps of 32 double-buffered in LDS with the next step
// prefetched into registers, v_mfma_f32_16x16x32_bf16 for the six products,
// FP32 partial sums flushed to FP64 every 64 values of k.  blockIdx.z picks
// one of ks slices of k (whole 64-chunks); slices above the first write
// their FP64 tile to a parts buffer, which k_add_parts sums in a fixed
// order.  The product is formed transposed (MFMA A operand from B) so each lane's outputs run down a
// column of C: coalesced stores.  The k order inside one MFMA is the same
// for both operands (lane group g holds k = 8 g .. 8 g + 7 of each), which is
// all the contraction needs.
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

constexpr int PL_BM = 128, PL_BN = 64, PL_BK = 32, PL_KCHUNK = 64;
constexpr int PL_NT = 512;  // eight waves per product tile
constexpr int PL_LDK = PL_BK + 8;  // shorts per LDS row (80 bytes)

template <bool KFAST>
__global__ __launch_bounds__(256) void k_split_planes(int R, int K, int rp, int kp, const double* __restrict__ x,
                                                      int ld, unsigned short* __restrict__ out) {
  __shared__ double tile[64][65];
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64, tid = threadIdx.x;
  for (int e = tid; e < 64 * 64; e += 256) {
    int r, k;
    if (KFAST) { r = e / 64; k = e % 64; } else { r = e % 64; k = e / 64; }
    const int gr = r0 + r, gk = k0 + k;
    double v = 0.0;
    if (gr < R && gk < K) v = KFAST ? x[(size_t)gk + (size_t)gr * ld] : x[(size_t)gr + (size_t)gk * ld];
    tile[r][k] = v;
  }
  __syncthreads();
  const size_t plane = (size_t)rp * kp;
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e / 64, k = e % 64;
    const int gr = r0 + r, gk = k0 + k;
    if (gr >= rp || gk >= kp) continue;
    unsigned short h0, h1, h2;
    split3(tile[r][k], h0, h1, h2);
    const size_t o = (size_t)gr * kp + gk;
    out[o] = h0;
    out[plane + o] = h1;
    out[2 * plane + o] = h2;
  }
}

__global__ __launch_bounds__(PL_NT) void k_gemm_split_planes(int M, int N, int kp, int mp, int np,
                                                           const unsigned short* __restrict__ ap,
                                                           const unsigned short* __restrict__ bp, double alpha,
                                                           double beta, double* __restrict__ c, int ldc,
                                                           int ks, double* __restrict__ cpart) {
  __shared__ __attribute__((aligned(16))) unsigned short As[2][3][PL_BM][PL_LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][3][PL_BN][PL_LDK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int m0 = blockIdx.y * PL_BM, n0 = blockIdx.x * PL_BN;
  // K slice blockIdx.z of ks (whole 64-chunks): the first adds beta C, the
  // others write alpha times their part to cpart (M x N each)
  const int kz = blockIdx.z, kslice = kp / ks, kbase = kz * kslice;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;  // eight waves, 32 x 32 each
  const size_t aplane = (size_t)mp * kp, bplane = (size_t)np * kp;
  constexpr int NA = 3 * PL_BM * PL_BK / 8 / PL_NT;  // 16-byte loads per thread: 3
  constexpr int NB_ALL = 3 * PL_BN * PL_BK / 8;     // 768 over the workgroup
  constexpr int NB = (NB_ALL + PL_NT - 1) / PL_NT;  // 2, the second for half the threads
  uint4 ra[NA], rb[NB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const int idx = tid + PL_NT * r;
      const int pl = idx / (PL_BM * 4), rem = idx % (PL_BM * 4);
      const int row = rem / 4, kc = (rem % 4) * 8;
      ra[r] = *reinterpret_cast<const uint4*>(ap + pl * aplane + (size_t)(m0 + row) * kp + kbase + k0 + kc);
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int idx = tid + PL_NT * r;
      if (idx >= NB_ALL) continue;
      const int pl = idx / (PL_BN * 4), rem = idx % (PL_BN * 4);
      const int row = rem / 4, kc = (rem % 4) * 8;
      rb[r] = *reinterpret_cast<const uint4*>(bp + pl * bplane + (size_t)(n0 + row) * kp + kbase + k0 + kc);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const int idx = tid + PL_NT * r;
      const int pl = idx / (PL_BM * 4), rem = idx % (PL_BM * 4);
      *reinterpret_cast<uint4*>(&As[buf][pl][rem / 4][(rem % 4) * 8]) = ra[r];
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int idx = tid + PL_NT * r;
      if (idx >= NB_ALL) continue;
      const int pl = idx / (PL_BN * 4), rem = idx % (PL_BN * 4);
      *reinterpret_cast<uint4*>(&Bs[buf][pl][rem / 4][(rem % 4) * 8]) = rb[r];
    }
  };
  f4_t big[2][2], corr[2][2];
  double tot[2][2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      big[j][i] = f4_t{0.f, 0.f, 0.f, 0.f};
      corr[j][i] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int t = 0; t < 4; ++t) tot[j][i][t] = 0.0;
    }
  gload(0);
  lstore(0);
  __syncthreads();
  const int nk = kslice / PL_BK;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * PL_BK);
    bf8_t af[2][3], bf[2][3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i][p] = *reinterpret_cast<const bf8_t*>(&As[cur][p][wm + 16 * i + li][8 * lg]);
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j][p] = *reinterpret_cast<const bf8_t*>(&Bs[cur][p][wn + 16 * j + li][8 * lg]);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        big[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j][0], af[i][0], big[j][i], 0, 0, 0);
        corr[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j][0], af[i][1], corr[j][i], 0, 0, 0);
        corr[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j][1], af[i][0], corr[j][i], 0, 0, 0);
        corr[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j][0], af[i][2], corr[j][i], 0, 0, 0);
        corr[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j][1], af[i][1], corr[j][i], 0, 0, 0);
        corr[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[j][2], af[i][0], corr[j][i], 0, 0, 0);
      }
    if (((kt + 1) * PL_BK) % PL_KCHUNK == 0 || kt + 1 == nk) {
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
          for (int t = 0; t < 4; ++t) tot[j][i][t] += (double)big[j][i][t] + (double)corr[j][i][t];
          big[j][i] = f4_t{0.f, 0.f, 0.f, 0.f};
          corr[j][i] = f4_t{0.f, 0.f, 0.f, 0.f};
        }
    }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  // lane holds C(m0 + wm + 16 i + li, n0 + wn + 16 j + 4 lg + t) in tot[j][i][t]
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gm = m0 + wm + 16 * i + li, gn = n0 + wn + 16 * j + 4 * lg + t;
        if (gm >= M || gn >= N) continue;
        const double v = tot[j][i][t];
        if (kz > 0) {
          cpart[(size_t)(kz - 1) * M * N + (size_t)gm + (size_t)gn * M] = alpha * v;
          continue;
        }
        double* dst = c + (size_t)gm + (size_t)gn * ldc;
        *dst = (beta == 0.0) ? alpha * v : fma(beta, *dst, alpha * v);
      }
}

// Plane scratch, one grow-only buffer per (device, stream): launches on one
// stream are ordered, so consecutive products reuse it safely.
struct PlaneScratch {
  int device;
  hipStream_t stream;
  DevBuf buf;
};
std::mutex g_plane_mu;
std::vector<PlaneScratch*> g_planes;

unsigned short* plane_scratch(hipStream_t s, size_t bytes) {
  int dev = 0;
  MHIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_plane_mu);
  for (PlaneScratch* p : g_planes)
    if (p->device == dev && p->stream == s) {
      p->buf.ensure(bytes);
      return p->buf.as<unsigned short>();
    }
  auto* p = new PlaneScratch{dev, s, DevBuf{}};
  g_planes.push_back(p);
  p->buf.ensure(bytes);
  return p->buf.as<unsigned short>();
}

}  // namespace

void release_split_scratch(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_plane_mu);
  std::vector<PlaneScratch*> keep;
  for (PlaneScratch* p : g_planes) {
    if (s == nullptr || p->stream == s) {
      p->buf.release();
      delete p;
    } else {
      keep.push_back(p);
    }
  }
  g_planes.swap(keep);
}

void gemm_split_bf16(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                     const double* a, int lda, const double* b, int ldb, double beta, double* c,
                     int ldc) {
  if (m <= 0 || n <= 0) return;
  const bool TA = (ta == 'T' || ta == 't'), TBT = (tb == 'T' || tb == 't');
  const int kp = cdiv_i(std::max(k, 1), PL_BK) * PL_BK;
  const int mp = cdiv_i(m, PL_BM) * PL_BM, np = cdiv_i(n, PL_BN) * PL_BN;
  const size_t abytes = (size_t)3 * mp * kp * sizeof(unsigned short);
  const size_t bbytes = (size_t)3 * np * kp * sizeof(unsigned short);
  unsigned short* planes = plane_scratch(s, abytes + bbytes);
  unsigned short* ap = planes;
  unsigned short* bp = planes + abytes / sizeof(unsigned short);
  timed(s, K_GEMM_SPLIT, 2.0 * m * n * (double)k, 8.0 * ((double)m * k + (double)k * n + 2.0 * m * n), [&] {
    // op(A)(m, k): A^T stored k-fast; op(B)(k, n) as rows n: B stored k-fast unless transposed
    const dim3 ga(cdiv_i(kp, 64), cdiv_i(mp, 64)), gb(cdiv_i(kp, 64), cdiv_i(np, 64));
    if (TA) hipLaunchKernelGGL(k_split_planes<true>, ga, dim3(256), 0, s, m, k, mp, kp, a, lda, ap);
    else hipLaunchKernelGGL(k_split_planes<false>, ga, dim3(256), 0, s, m, k, mp, kp, a, lda, ap);
    MHIP_LAUNCH_CHECK();
    if (!TBT) hipLaunchKernelGGL(k_split_planes<true>, gb, dim3(256), 0, s, n, k, np, kp, b, ldb, bp);
    else hipLaunchKernelGGL(k_split_planes<false>, gb, dim3(256), 0, s, n, k, np, kp, b, ldb, bp);
    MHIP_LAUNCH_CHECK();
    // one workgroup per CU (LDS): split K (whole 64-chunks, 256 or more per
    // slice) when it balances the CUs better
    const long long wgs = (long long)(np / PL_BN) * (mp / PL_BM);
    int ks = 1;
    double ks_cost = (double)((wgs + 255) / 256);
    for (int q : {2, 4}) {
      if (kp % (q * PL_KCHUNK) != 0 || kp / q < 256) continue;
      const double cost = (double)((wgs * q + 255) / 256) / q;
      if (cost < ks_cost - 1e-9) {
        ks = q;
        ks_cost = cost;
      }
    }
    double* cp = ks > 1 ? splitk_scratch(s, sizeof(double) * (size_t)m * n * (ks - 1)) : nullptr;
    hipLaunchKernelGGL(k_gemm_split_planes, dim3(np / PL_BN, mp / PL_BM, ks), dim3(PL_NT), 0, s, m, n, kp, mp, np,
                       ap, bp, alpha, beta, c, ldc, ks, cp);
    MHIP_LAUNCH_CHECK();
    if (ks > 1) add_parts(s, m, n, ks - 1, cp, c, ldc);
  });
}

}  // namespace mhip
