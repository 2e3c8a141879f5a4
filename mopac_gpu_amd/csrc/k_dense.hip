// Dense FP64 building blocks: tiled GEMM, SYRK into MOPAC packed storage,
// packed <-> dense conversion, the packed Pulay commutator, linear
// combinations and deterministic dot products.
#include <cstdint>

#include "common.h"
#include <mutex>
#include <vector>
#include "mfma.h"

namespace mhip {

KTimer g_ktimer;
namespace {

constexpr int TB = 64;    // output tile edge
constexpr int TK = 16;    // k chunk
constexpr int NT = 256;   // threads per workgroup (16 x 16, 4x4 outputs each)

// op(A)(m,k), op(B)(k,n) loaders for column-major storage.
template <bool TA>
__device__ __forceinline__ double ld_a(const double* a, int lda, int m, int k, int M, int K) {
  if (m >= M || k >= K) return 0.0;
  return TA ? a[(size_t)k + (size_t)m * lda] : a[(size_t)m + (size_t)k * lda];
}
template <bool TBT>
__device__ __forceinline__ double ld_b(const double* b, int ldb, int k, int n, int K, int N) {
  if (k >= K || n >= N) return 0.0;
  return TBT ? b[(size_t)n + (size_t)k * ldb] : b[(size_t)k + (size_t)n * ldb];
}

// One 64x64 tile of op(A) op(B); acc[i][j] is row ty*4+i, col tx*4+j.
template <bool TA, bool TBT>
__device__ __forceinline__ void tile_mm(const double* a, int lda, const double* b, int ldb, int M,
                                        int N, int K, int row0, int col0, double acc[4][4],
                                        double (*As)[TB + 1], double (*Bs)[TB + 1]) {
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  for (int k0 = 0; k0 < K; k0 += TK) {
    // A tile: 64 (m) x 16 (k)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * NT;
      int mm, kk;
      if (TA) { kk = e & 15; mm = e >> 4; } else { mm = e & 63; kk = e >> 6; }
      As[kk][mm] = ld_a<TA>(a, lda, row0 + mm, k0 + kk, M, K);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * NT;
      int kk, nn;
      if (TBT) { nn = e & 63; kk = e >> 6; } else { kk = e & 15; nn = e >> 4; }
      Bs[kk][nn] = ld_b<TBT>(b, ldb, k0 + kk, col0 + nn, K, N);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
}

template <bool TA, bool TBT>
__global__ __launch_bounds__(NT) void k_gemm(int M, int N, int K, double alpha, const double* a,
                                             int lda, const double* b, int ldb, double beta,
                                             double* c, int ldc) {
  __shared__ double As[TK][TB + 1];
  __shared__ double Bs[TK][TB + 1];
  const int row0 = blockIdx.y * TB, col0 = blockIdx.x * TB;
  double acc[4][4] = {};
  tile_mm<TA, TBT>(a, lda, b, ldb, M, N, K, row0, col0, acc, As, Bs);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = row0 + ty * 4 + i;
    if (r >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int cc = col0 + tx * 4 + j;
      if (cc >= N) continue;
      double* dst = c + (size_t)r + (size_t)cc * ldc;
      const double v = alpha * acc[i][j];
      *dst = (beta == 0.0) ? v : fma(beta, *dst, v);
    }
  }
}

// Lower-triangle tiles of occ * A A^T written straight into packed storage.
__global__ __launch_bounds__(NT) void k_syrk_packed(int n, int k, double occ, const double* a,
                                                    int lda, double* p) {
  __shared__ double As[TK][TB + 1];
  __shared__ double Bs[TK][TB + 1];
  // blockIdx.x enumerates (bi >= bj) tiles
  const int t = blockIdx.x;
  int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  while (bi * (bi + 1) / 2 > t) --bi;
  const int bj = t - bi * (bi + 1) / 2;
  const int row0 = bi * TB, col0 = bj * TB;
  double acc[4][4] = {};
  tile_mm<false, true>(a, lda, a, lda, n, n, k, row0, col0, acc, As, Bs);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long long r = row0 + ty * 4 + i;
    if (r >= n) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long cc = col0 + tx * 4 + j;
      if (cc > r) continue;
      p[r * (r + 1) / 2 + cc] = occ * acc[i][j];
    }
  }
}

__global__ void k_unpack_sym(int n, const double* __restrict__ p, double* __restrict__ a, int lda,
                             int npad, double pad_base) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= npad || j >= npad) return;
  double v;
  if (i < n && j < n) {
    const long long r = i > j ? i : j, c = i > j ? j : i;
    v = p[r * (r + 1) / 2 + c];
  } else {
    v = (i == j) ? pad_base + (double)i : 0.0;
  }
  a[(size_t)i + (size_t)j * lda] = v;
}

__global__ void k_pack_lower(int n, const double* __restrict__ a, int lda, double* __restrict__ p) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (r >= n || c > r) return;
  p[(long long)r * (r + 1) / 2 + c] = a[(size_t)r + (size_t)c * lda];
}

__global__ void k_pack_antisym(int n, const double* __restrict__ x, int ldx,
                               double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (r >= n || c > r) return;
  out[(long long)r * (r + 1) / 2 + c] = x[(size_t)c + (size_t)r * ldx] - x[(size_t)r + (size_t)c * ldx];
}

__global__ void k_lincomb(long long len, int nk, const double* __restrict__ coef,
                          const double* __restrict__ x, long long stride, double* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len;
       i += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < nk; ++k) s += coef[k] * x[(size_t)k * stride + i];
    y[i] = s;
  }
}

__global__ void k_fill(double* p, long long n, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

constexpr int DOT_BLOCKS = 512;

__global__ __launch_bounds__(256) void k_dots_partial(long long len, int nvec, const double* a,
                                                      long long stride, const double* b,
                                                      double* partial) {
  __shared__ double red[256];
  const int v = blockIdx.y;
  const double* av = a + (size_t)v * stride;
  double s = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < len; i += (long long)gridDim.x * 256)
    s = fma(av[i], b[i], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)v * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_dots_final(int nblocks, const double* partial,
                                                    double* out) {
  __shared__ double red[256];
  const int v = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 256) s += partial[(size_t)v * nblocks + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = red[0];
}


// Same tiling on the FP64 matrix cores (see mfma.h for the lane layout).
template <bool TA, bool TBT>
__global__ __launch_bounds__(NT) void k_gemm_mfma(int M, int N, int K, double alpha,
                                                  const double* a, int lda, const double* b,
                                                  int ldb, double beta, double* c, int ldc) {
  __shared__ double As[TK][TB + 1];
  __shared__ double Bs[TK][TB + 1];
  const int row0 = blockIdx.y * TB, col0 = blockIdx.x * TB;
  const int tid = threadIdx.x;
  d4_t acc[4];
  for (int i = 0; i < 4; ++i) acc[i] = d4_t{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += TK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * NT;
      int mm, kk;
      if (TA) { kk = e & 15; mm = e >> 4; } else { mm = e & 63; kk = e >> 6; }
      As[kk][mm] = ld_a<TA>(a, lda, row0 + mm, k0 + kk, M, K);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * NT;
      int kk, nn;
      if (TBT) { nn = e & 63; kk = e >> 6; } else { kk = e & 15; nn = e >> 4; }
      Bs[kk][nn] = ld_b<TBT>(b, ldb, k0 + kk, col0 + nn, K, N);
    }
    __syncthreads();
    mfma_tile64([&](int r, int k) { return As[k][r]; }, [&](int k, int cc) { return Bs[k][cc]; },
                TK, acc);
    __syncthreads();
  }
  mfma_tile64_store(acc, [&](int r, int cc, double v) {
    const int gr = row0 + r, gc = col0 + cc;
    if (gr >= M || gc >= N) return;
    double* dst = c + (size_t)gr + (size_t)gc * ldc;
    *dst = (beta == 0.0) ? alpha * v : fma(beta, *dst, alpha * v);
  });
}

// Large-tile FP64 MFMA GEMM for shapes that divide the tile: BM x BN per
// workgroup, four waves in 2 x 2 each owning (BM/2) x (BN/2) as 16 x 16 MFMA
// blocks.  k advances in chunks of 16 that are double-buffered in LDS; the
// next chunk is fetched from global memory into registers (16-byte loads)
// while the current one is multiplied, so one barrier per chunk.  The product
// is formed transposed (the MFMA A operand is taken from B, the B operand
// from A), which puts each lane's outputs down a column of C: stores are
// coalesced along the column-major rows.  LDS rows are padded to a stride of
// 16 (mod 32) doubles so a half-wave's 16-lane groups hit disjoint banks.
// Requires M % BM == 0, N % BN == 0, K % 16 == 0, K > 0, 16-byte aligned a, b
// and even lda, ldb (checked by the launcher).
// Workgroup -> output tile.  Workgroups are dealt round-robin to the eight
// XCDs, each with its own L2: the first step gives every XCD a contiguous
// chunk of the tile sequence; the second orders that sequence in groups of
// eight tile rows, column by column, so one XCD's chunk is a compact block
// that shares A row panels and B column panels in its L2.
__device__ __forceinline__ void xcd_tile(int L, int gm, int gn, int& tm, int& tn) {
  const int T = gm * gn;
  const int per = T / 8;
  const int t = (L < per * 8) ? (L % 8) * per + L / 8 : L;
  constexpr int G = 8;
  const int first = (t / (G * gn)) * G;
  const int rows = min(gm - first, G);
  const int r = t - first * gn;
  tm = first + r % rows;
  tn = r / rows;
}

// Lower-triangle tile sequence (tile column by tile column, rows tn..T-1),
// dealt to the XCDs in contiguous chunks as in xcd_tile, so one XCD holds a
// few whole tile columns that share their B panel.
__device__ __forceinline__ void tri_tile(int L, int ntile, int T, int& tm, int& tn) {
  const int per = ntile / 8;
  int t = (L < per * 8) ? (L % 8) * per + L / 8 : L;
  int col = 0;
  while (t >= T - col) {
    t -= T - col;
    ++col;
  }
  tn = col;
  tm = col + t;
}

constexpr int BK = 16;
// Eight waves (two per SIMD) on a 4 x 2 grid of sub-tiles, so each SIMD has
// a second wave to issue MFMAs while one waits on LDS or the barrier.
constexpr int GB_NT = 512;
template <int BM, int BN, bool TA, bool TBT>
__global__ __launch_bounds__(GB_NT) void k_gemm_big(int M, int N, int K, double alpha,
                                                    const double* __restrict__ a, int lda,
                                                    const double* __restrict__ b, int ldb, double beta,
                                                    double* __restrict__ c, int ldc, int ksplit,
                                                    double* __restrict__ cpart, int tri) {
  constexpr int SA = BM + 16, SBn = BN + 16;
  constexpr int WM = BM / 4, WN = BN / 2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int NA2 = BM * BK / 2 / GB_NT, NB2 = BN * BK / 2 / GB_NT;
  static_assert(NA2 >= 1 && NB2 >= 1 && FM >= 1 && FN >= 1, "tile too small");
  __shared__ double As[2][BK][SA];
  __shared__ double Bs[2][BK][SBn];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  int tm, tn;
  // ksplit > 1: workgroup kz of a tile takes the kz-th K / ksplit slice; the
  // first adds beta C, the others write alpha times their part to cpart
  // (M x N per extra slice), summed into C by k_add_parts
  // tri (M == N, BM == BN): only the tiles on and below the diagonal
  const int gtm = M / BM, gtn = N / BN;
  const int ntile = tri ? gtm * (gtm + 1) / 2 : gtm * gtn;
  const int kz = blockIdx.x / ntile;
  const int kbase = kz * (K / ksplit);
  if (tri) tri_tile(blockIdx.x - kz * ntile, ntile, gtm, tm, tn);
  else xcd_tile(blockIdx.x - kz * ntile, gtm, gtn, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int wm = (w >> 1) * WM, wn = (w & 1) * WN;  // w = 0..7
  // operands stored transposed (A when TA, B when !TBT): a wave writes 8 k
  // rows of one column group, rows that share banks; the column index is
  // XOR-ed with the k pair (bits 2-4), which spreads those stores over the
  // banks and keeps every aligned 16-column block together for the reads
  auto swz = [](int k) { return ((k >> 1) & 7) << 2; };
  double2 ra[NA2], rb[NB2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NA2; ++r) {
      const int idx = tid + GB_NT * r;
      const double* src;
      if (!TA) src = a + (size_t)(m0 + 2 * (idx % (BM / 2))) + (size_t)(kbase + k0 + idx / (BM / 2)) * lda;
      else src = a + (size_t)(kbase + k0 + 2 * (idx % (BK / 2))) + (size_t)(m0 + idx / (BK / 2)) * lda;
      ra[r] = *reinterpret_cast<const double2*>(src);
    }
#pragma unroll
    for (int r = 0; r < NB2; ++r) {
      const int idx = tid + GB_NT * r;
      const double* src;
      if (!TBT) src = b + (size_t)(kbase + k0 + 2 * (idx % (BK / 2))) + (size_t)(n0 + idx / (BK / 2)) * ldb;
      else src = b + (size_t)(n0 + 2 * (idx % (BN / 2))) + (size_t)(kbase + k0 + idx / (BN / 2)) * ldb;
      rb[r] = *reinterpret_cast<const double2*>(src);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < NA2; ++r) {
      const int idx = tid + GB_NT * r;
      if (!TA) {
        const int mm = 2 * (idx % (BM / 2)), kk = idx / (BM / 2);
        As[buf][kk][mm] = ra[r].x;
        As[buf][kk][mm + 1] = ra[r].y;
      } else {
        const int kk = 2 * (idx % (BK / 2)), mm = idx / (BK / 2);
        As[buf][kk][mm ^ swz(kk)] = ra[r].x;
        As[buf][kk + 1][mm ^ swz(kk + 1)] = ra[r].y;
      }
    }
#pragma unroll
    for (int r = 0; r < NB2; ++r) {
      const int idx = tid + GB_NT * r;
      if (!TBT) {
        const int kk = 2 * (idx % (BK / 2)), nn = idx / (BK / 2);
        Bs[buf][kk][nn ^ swz(kk)] = rb[r].x;
        Bs[buf][kk + 1][nn ^ swz(kk + 1)] = rb[r].y;
      } else {
        const int nn = 2 * (idx % (BN / 2)), kk = idx / (BN / 2);
        Bs[buf][kk][nn] = rb[r].x;
        Bs[buf][kk][nn + 1] = rb[r].y;
      }
    }
  };
  d4_t acc[FN][FM];
#pragma unroll
  for (int j = 0; j < FN; ++j)
#pragma unroll
    for (int i = 0; i < FM; ++i) acc[j][i] = d4_t{0.0, 0.0, 0.0, 0.0};
  gload(0);
  lstore(0);
  __syncthreads();
  const int nk = K / ksplit / BK;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[cur][ks + lk][(wm + 16 * i + li) ^ (TA ? swz(ks + lk) : 0)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[cur][ks + lk][(wn + 16 * j + li) ^ (TBT ? 0 : swz(ks + lk))];
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int i = 0; i < FM; ++i) acc[j][i] = mfma16x16x4(bf[j], af[i], acc[j][i]);
    }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  // lane holds C(m0 + wm + 16 i + li, n0 + wn + 16 j + lk + 4 t) in acc[j][i][t]
#pragma unroll
  for (int j = 0; j < FN; ++j)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gm = m0 + wm + 16 * i + li, gn = n0 + wn + 16 * j + lk + 4 * t;
        const double v = acc[j][i][t];
        if (kz > 0) {
          cpart[(size_t)(kz - 1) * M * N + (size_t)gm + (size_t)gn * M] = alpha * v;
          continue;
        }
        double* dst = c + (size_t)gm + (size_t)gn * ldc;
        *dst = (beta == 0.0) ? alpha * v : fma(beta, *dst, alpha * v);
      }
}

// Packed SYRK on the large tiles: p = occ * A A^T over the lower triangle
// of 128 x 128 tiles (XCD-chunked), the k loop of k_gemm_big with both
// operands taken from rows of A (column-major, n x k).  The MFMA output keeps
// lanes along columns, which are contiguous in the packed rows.  Requires
// n % 128 == 0, k % 16 == 0, k > 0, 16-byte aligned a and even lda.
__global__ __launch_bounds__(GB_NT) void k_syrk_big(int n, int k, double occ, const double* __restrict__ a,
                                                    int lda, double* __restrict__ p, int ksplit,
                                                    double* __restrict__ ppart) {
  // eight waves on a 4 x 2 grid of 32 x 64 sub-tiles, as in k_gemm_big
  constexpr int BM = 128, S = BM + 16, WM = 32, WN = 64, FM = 2, FN = 4, NL = BM * BK / 2 / GB_NT;
  __shared__ double As[2][BK][S];
  __shared__ double Bs[2][BK][S];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nt = n / BM, T = nt * (nt + 1) / 2;
  // ksplit > 1: workgroup kz of a tile takes the kz-th K / ksplit slice; the
  // first writes p, the others write their packed part to ppart
  const int kz = blockIdx.x / T, kbase = kz * (k / ksplit);
  const int per = T / 8, L = blockIdx.x - kz * T;
  const int t = (L < per * 8) ? (L % 8) * per + L / 8 : L;
  int tm = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((tm + 1) * (tm + 2) / 2 <= t) ++tm;
  while (tm * (tm + 1) / 2 > t) --tm;
  const int tn = t - tm * (tm + 1) / 2;
  const int m0 = tm * BM, n0 = tn * BM;
  const int wm = (w >> 1) * WM, wn = (w & 1) * WN;
  double2 ra[NL], rb[NL];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = tid + GB_NT * r;
      const int mm = 2 * (idx % (BM / 2)), kk = idx / (BM / 2);
      ra[r] = *reinterpret_cast<const double2*>(a + (size_t)(m0 + mm) + (size_t)(kbase + k0 + kk) * lda);
      rb[r] = *reinterpret_cast<const double2*>(a + (size_t)(n0 + mm) + (size_t)(kbase + k0 + kk) * lda);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int idx = tid + GB_NT * r;
      const int mm = 2 * (idx % (BM / 2)), kk = idx / (BM / 2);
      As[buf][kk][mm] = ra[r].x;
      As[buf][kk][mm + 1] = ra[r].y;
      Bs[buf][kk][mm] = rb[r].x;
      Bs[buf][kk][mm + 1] = rb[r].y;
    }
  };
  d4_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
  gload(0);
  lstore(0);
  __syncthreads();
  const int nk = k / ksplit / BK;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[cur][ks + lk][wm + 16 * i + li];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[cur][ks + lk][wn + 16 * j + li];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = mfma16x16x4(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }
  // lane holds (m0 + wm + 16 i + lk + 4 t, n0 + wn + 16 j + li) in acc[i][j][t]
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long long gm = m0 + wm + 16 * i + lk + 4 * q, gn = n0 + wn + 16 * j + li;
        if (gn > gm) continue;
        if (kz > 0) ppart[(size_t)(kz - 1) * ((size_t)n * (n + 1) / 2) + gm * (gm + 1) / 2 + gn] = occ * acc[i][j][q];
        else p[gm * (gm + 1) / 2 + gn] = occ * acc[i][j][q];
      }
}

__global__ __launch_bounds__(NT) void k_syrk_packed_mfma(int n, int k, double occ, const double* a,
                                                         int lda, double* p) {
  __shared__ double As[TK][TB + 1];
  __shared__ double Bs[TK][TB + 1];
  const int t = blockIdx.x;
  int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  while (bi * (bi + 1) / 2 > t) --bi;
  const int bj = t - bi * (bi + 1) / 2;
  const int row0 = bi * TB, col0 = bj * TB;
  const int tid = threadIdx.x;
  d4_t acc[4];
  for (int i = 0; i < 4; ++i) acc[i] = d4_t{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < k; k0 += TK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + r * NT;
      const int mm = e & 63, kk = e >> 6;
      As[kk][mm] = ld_a<false>(a, lda, row0 + mm, k0 + kk, n, k);
      Bs[kk][mm] = ld_a<false>(a, lda, col0 + mm, k0 + kk, n, k);
    }
    __syncthreads();
    mfma_tile64([&](int r, int kk) { return As[kk][r]; }, [&](int kk, int cc) { return Bs[kk][cc]; },
                TK, acc);
    __syncthreads();
  }
  mfma_tile64_store(acc, [&](int r, int cc, double v) {
    const long long gr = row0 + r, gc = col0 + cc;
    if (gr >= n || gc > gr) return;
    p[gr * (gr + 1) / 2 + gc] = occ * v;
  });
}

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace

int g_kernel_variant = 1;  // 1: FP64 MFMA tiles, 0: vector-FMA tiles

template <int BM, int BN>
static void gemm_big_dispatch(hipStream_t s, bool TA, bool TBT, int m, int n, int k, double alpha,
                              const double* a, int lda, const double* b, int ldb, double beta, double* c,
                              int ldc, int ks, double* cp, int tri = 0) {
  const int tiles = tri ? (m / BM) * (m / BM + 1) / 2 : (m / BM) * (n / BN);
  const dim3 grid(tiles * ks);
  if (!TA && !TBT) hipLaunchKernelGGL((k_gemm_big<BM, BN, false, false>), grid, dim3(GB_NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp, tri);
  else if (TA && !TBT) hipLaunchKernelGGL((k_gemm_big<BM, BN, true, false>), grid, dim3(GB_NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp, tri);
  else if (!TA && TBT) hipLaunchKernelGGL((k_gemm_big<BM, BN, false, true>), grid, dim3(GB_NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp, tri);
  else hipLaunchKernelGGL((k_gemm_big<BM, BN, true, true>), grid, dim3(GB_NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp, tri);
  MHIP_LAUNCH_CHECK();
}

// C += the K-slice parts the split workgroups wrote (M x N each, in order)
__global__ void k_add_parts(int m, int n, int nparts, const double* __restrict__ part, double* __restrict__ c,
                            int ldc) {
  const size_t mn = (size_t)m * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % m), j = (int)(e / m);
    double v = c[i + (size_t)j * ldc];
    for (int q = 0; q < nparts; ++q) v += part[q * mn + e];
    c[i + (size_t)j * ldc] = v;
  }
}

void add_parts(hipStream_t s, int m, int n, int nparts, const double* part, double* c, int ldc) {
  hipLaunchKernelGGL(k_add_parts, dim3(2048), dim3(256), 0, s, m, n, nparts, part, c, ldc);
  MHIP_LAUNCH_CHECK();
}

// K-slice scratch, one grow-only buffer per (device, stream): launches on one
// stream are ordered, so consecutive products reuse it safely.
struct SplitKScratch {
  int device;
  hipStream_t stream;
  DevBuf buf;
};
std::mutex g_splitk_mu;
std::vector<SplitKScratch*> g_splitk;

double* splitk_scratch(hipStream_t s, size_t bytes) {
  int dev = 0;
  MHIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_splitk_mu);
  for (SplitKScratch* p : g_splitk)
    if (p->device == dev && p->stream == s) {
      p->buf.ensure(bytes);
      return p->buf.as<double>();
    }
  auto* p = new SplitKScratch{dev, s, DevBuf{}};
  g_splitk.push_back(p);
  p->buf.ensure(bytes);
  return p->buf.as<double>();
}

// The large-tile kernel when the shape divides one of its tiles; the tile is
// the one with the least estimated time: rounds of 256 workgroups (one per
// CU) times the tile's work over its relative efficiency.
static bool gemm_big_launch(hipStream_t s, bool TA, bool TBT, int m, int n, int k, double alpha,
                            const double* a, int lda, const double* b, int ldb, double beta, double* c,
                            int ldc) {
  if (k < BK || k % BK != 0 || (lda & 1) || (ldb & 1)) return false;
  if ((reinterpret_cast<uintptr_t>(a) & 15) || (reinterpret_cast<uintptr_t>(b) & 15)) return false;
  struct Cand { int bm, bn; double eff; };
  const Cand cands[3] = {{128, 128, 1.0}, {128, 64, 0.85}, {64, 64, 0.6}};
  int best = -1;
  double best_cost = 0.0;
  for (int i = 0; i < 3; ++i) {
    const Cand& q = cands[i];
    if (m % q.bm || n % q.bn) continue;
    const long long tiles = (long long)(m / q.bm) * (n / q.bn);
    const double cost = (double)((tiles + 255) / 256) * q.bm * q.bn / q.eff;
    if (best < 0 || cost < best_cost) {
      best = i;
      best_cost = cost;
    }
  }
  if (best < 0) return false;
  // K split in 2, 3 or 4 (each slice keeps 256 or more of K) when it balances the CUs
  // better: the fewest full-tile equivalents per CU, ceil(tiles ks / 256) / ks,
  // ties to the smaller split; one-round launches of the smaller tiles always
  // split in two (their workgroups gain most from a second one on the CU)
  const long long tiles = (long long)(m / cands[best].bm) * (n / cands[best].bn);
  int ks = 1;
  double ks_cost = (double)((tiles + 255) / 256);
  for (int q : {2, 3, 4}) {
    if (k % (q * BK) != 0 || k / q < 256) continue;
    const double cost = (double)((tiles * q + 255) / 256) / q;
    if (cost < ks_cost - 1e-9 || (q == 2 && best != 0 && tiles <= 256)) {
      ks = q;
      ks_cost = cost;
    }
  }
  double* cp = ks > 1 ? splitk_scratch(s, sizeof(double) * (size_t)m * n * (ks - 1)) : nullptr;
  if (best == 0) gemm_big_dispatch<128, 128>(s, TA, TBT, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp);
  else if (best == 1) gemm_big_dispatch<128, 64>(s, TA, TBT, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp);
  else gemm_big_dispatch<64, 64>(s, TA, TBT, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp);
  if (ks > 1) add_parts(s, m, n, ks - 1, cp, c, ldc);
  return true;
}

static void gemm_launch(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                        const double* a, int lda, const double* b, int ldb, double beta, double* c,
                        int ldc) {
  if (m <= 0 || n <= 0) return;
  dim3 grid(cdiv(n, TB), cdiv(m, TB));
  const bool TA = (ta == 'T' || ta == 't'), TBT = (tb == 'T' || tb == 't');
  if (g_kernel_variant == 1 && gemm_big_launch(s, TA, TBT, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc))
    return;
  if (g_kernel_variant == 1) {
    if (!TA && !TBT) hipLaunchKernelGGL((k_gemm_mfma<false, false>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    else if (TA && !TBT) hipLaunchKernelGGL((k_gemm_mfma<true, false>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    else if (!TA && TBT) hipLaunchKernelGGL((k_gemm_mfma<false, true>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    else hipLaunchKernelGGL((k_gemm_mfma<true, true>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
    MHIP_LAUNCH_CHECK();
    return;
  }
  if (!TA && !TBT) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
  else if (TA && !TBT) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
  else if (!TA && TBT) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
  else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
  MHIP_LAUNCH_CHECK();
}

void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double* a,
          int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  if (m <= 0 || n <= 0) return;
  timed(s, K_GEMM_F64, 2.0 * m * n * (double)k, 8.0 * ((double)m * k + (double)k * n + 2.0 * m * n),
        [&] { gemm_launch(s, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}

// Symmetric m x m result from its lower-triangle tiles: each workgroup takes
// one 32 x 32 block (bi, bj), bi >= bj, adds the K-slice parts (m x m each,
// written on the lower tiles only), and stores the block and its transpose,
// both with lanes down columns.  A diagonal block takes its upper half from
// its own lower half, so the result is exactly symmetric.
__global__ __launch_bounds__(256) void k_sym_from_lower(int m, int nparts, const double* __restrict__ part,
                                                        double* __restrict__ c, int ldc) {
  __shared__ double t[32][33];
  int b = blockIdx.x, bj = 0;
  const int nb = m / 32;
  while (b >= nb - bj) {
    b -= nb - bj;
    ++bj;
  }
  const int bi = bj + b;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const size_t mm = (size_t)m * m;
  for (int r = ty; r < 32; r += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + r;
    double v = c[i + (size_t)j * ldc];
    for (int q = 0; q < nparts; ++q) v += part[q * mm + i + (size_t)j * m];
    t[tx][r] = v;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = bi * 32 + tx, j = bj * 32 + r;
    c[i + (size_t)j * ldc] = (bi == bj && tx < r) ? t[r][tx] : t[tx][r];
    if (bi != bj) c[(size_t)(bj * 32 + tx) + (size_t)(bi * 32 + r) * ldc] = t[r][tx];
  }
}

// Shapes and operands the lower-triangle launch takes.
static bool gemm_sym_fits(int m, int k, const double* a, int lda, const double* b, int ldb) {
  if (g_kernel_variant != 1 || m % 64 != 0 || k < BK || k % BK != 0 || (lda & 1) || (ldb & 1)) return false;
  return !(reinterpret_cast<uintptr_t>(a) & 15) && !(reinterpret_cast<uintptr_t>(b) & 15);
}

// The large-tile kernel on the lower-triangle tiles of a square result.
static void gemm_sym_launch(hipStream_t s, bool TA, bool TBT, int m, int k, double alpha, const double* a,
                            int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  // square tiles only; estimated time as in gemm_big_launch over the triangle
  int bm = 0;
  double best_cost = 0.0;
  for (const auto& q : {std::pair<int, double>{128, 1.0}, std::pair<int, double>{64, 0.6}}) {
    if (m % q.first) continue;
    const long long T = m / q.first, tiles = T * (T + 1) / 2;
    const double cost = (double)((tiles + 255) / 256) * q.first * q.first / q.second;
    if (bm == 0 || cost < best_cost) {
      bm = q.first;
      best_cost = cost;
    }
  }
  const long long T = m / bm, tiles = T * (T + 1) / 2;
  int ks = 1;
  double ks_cost = (double)((tiles + 255) / 256);
  for (int q : {2, 3, 4}) {
    if (k % (q * BK) != 0 || k / q < 256) continue;
    const double cost = (double)((tiles * q + 255) / 256) / q;
    if (cost < ks_cost - 1e-9 || (q == 2 && bm != 128 && tiles <= 256)) {
      ks = q;
      ks_cost = cost;
    }
  }
  double* cp = ks > 1 ? splitk_scratch(s, sizeof(double) * (size_t)m * m * (ks - 1)) : nullptr;
  if (bm == 128) gemm_big_dispatch<128, 128>(s, TA, TBT, m, m, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp, 1);
  else gemm_big_dispatch<64, 64>(s, TA, TBT, m, m, k, alpha, a, lda, b, ldb, beta, c, ldc, ks, cp, 1);
  const int nb = m / 32;
  hipLaunchKernelGGL(k_sym_from_lower, dim3(nb * (nb + 1) / 2), dim3(256), 0, s, m, ks - 1, cp, c, ldc);
  MHIP_LAUNCH_CHECK();
}

void gemm_sym(hipStream_t s, char ta, char tb, int m, int k, double alpha, const double* a, int lda,
              const double* b, int ldb, double beta, double* c, int ldc) {
  if (m <= 0) return;
  if (!gemm_sym_fits(m, k, a, lda, b, ldb)) {
    gemm(s, ta, tb, m, m, k, alpha, a, lda, b, ldb, beta, c, ldc);
    return;
  }
  const bool TA = (ta == 'T' || ta == 't'), TBT = (tb == 'T' || tb == 't');
  timed(s, K_GEMM_F64, (double)m * (m + 1) * (double)k, 8.0 * (2.0 * m * k + (double)m * m),
        [&] { gemm_sym_launch(s, TA, TBT, m, k, alpha, a, lda, b, ldb, beta, c, ldc); });
}

static void syrk_launch(hipStream_t s, int n, int k, double occ, const double* a, int lda, double* p) {
  if (g_kernel_variant == 1 && n % 128 == 0 && k >= BK && k % BK == 0 && !(lda & 1) &&
      !(reinterpret_cast<uintptr_t>(a) & 15)) {
    const int nb = n / 128;
    // K split for CU balance, as in gemm_big_launch (packed parts)
    const long long T = (long long)nb * (nb + 1) / 2;
    int ks = 1;
    double ks_cost = (double)((T + 255) / 256);
    for (int q : {2, 4}) {
      if (k % (q * BK) != 0 || k / q < 256) continue;
      const double cost = (double)((T * q + 255) / 256) / q;
      if (cost < ks_cost - 1e-9) {
        ks = q;
        ks_cost = cost;
      }
    }
    const size_t mp = (size_t)n * (n + 1) / 2;
    double* cp = ks > 1 ? splitk_scratch(s, sizeof(double) * mp * (ks - 1)) : nullptr;
    hipLaunchKernelGGL(k_syrk_big, dim3(T * ks), dim3(GB_NT), 0, s, n, k, occ, a, lda, p, ks, cp);
    MHIP_LAUNCH_CHECK();
    if (ks > 1) add_parts(s, (int)mp, 1, ks - 1, cp, p, (int)mp);
    return;
  }
  const int nt = cdiv(n, TB);
  if (g_kernel_variant == 1)
    hipLaunchKernelGGL(k_syrk_packed_mfma, dim3(nt * (nt + 1) / 2), dim3(NT), 0, s, n, k, occ, a, lda, p);
  else
    hipLaunchKernelGGL(k_syrk_packed, dim3(nt * (nt + 1) / 2), dim3(NT), 0, s, n, k, occ, a, lda, p);
  MHIP_LAUNCH_CHECK();
}

void syrk_packed(hipStream_t s, int n, int k, double occ, const double* a, int lda, double* p) {
  if (n <= 0) return;
  timed(s, K_SYRK_F64, (double)n * (n + 1) * (double)k, 8.0 * ((double)n * k + (double)n * (n + 1) / 2),
        [&] { syrk_launch(s, n, k, occ, a, lda, p); });
}

void unpack_sym(hipStream_t s, int n, const double* p, double* a, int lda, int npad,
                double pad_base) {
  hipLaunchKernelGGL(k_unpack_sym, dim3(cdiv(npad, 256), npad), dim3(256), 0, s, n, p, a, lda, npad,
                     pad_base);
  MHIP_LAUNCH_CHECK();
}

void pack_lower(hipStream_t s, int n, const double* a, int lda, double* p) {
  hipLaunchKernelGGL(k_pack_lower, dim3(cdiv(n, 256), n), dim3(256), 0, s, n, a, lda, p);
  MHIP_LAUNCH_CHECK();
}

void pack_antisym(hipStream_t s, int n, const double* x, int ldx, double* out) {
  hipLaunchKernelGGL(k_pack_antisym, dim3(cdiv(n, 256), n), dim3(256), 0, s, n, x, ldx, out);
  MHIP_LAUNCH_CHECK();
}

void lincomb(hipStream_t s, long long len, int nk, const double* coef_dev, const double* x,
             long long stride, double* y) {
  const int blocks = std::min(4096, cdiv(len, 256));
  hipLaunchKernelGGL(k_lincomb, dim3(blocks), dim3(256), 0, s, len, nk, coef_dev, x, stride, y);
  MHIP_LAUNCH_CHECK();
}

void fill(hipStream_t s, double* p, long long n, double v) {
  if (n <= 0) return;
  const int blocks = std::min(4096, cdiv(n, 256));
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, n, v);
  MHIP_LAUNCH_CHECK();
}

size_t dots_partial_doubles(long long len, int nvec) {
  (void)len;
  return (size_t)DOT_BLOCKS * (size_t)nvec;
}

void dots(hipStream_t s, long long len, int nvec, const double* a, long long stride,
          const double* b, double* out, double* partial) {
  if (nvec <= 0) return;
  hipLaunchKernelGGL(k_dots_partial, dim3(DOT_BLOCKS, nvec), dim3(256), 0, s, len, nvec, a, stride,
                     b, partial);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dots_final, dim3(nvec), dim3(256), 0, s, DOT_BLOCKS, partial, out);
  MHIP_LAUNCH_CHECK();
}

}  // namespace mhip

namespace mhip {
void release_splitk_scratch(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_splitk_mu);
  std::vector<SplitKScratch*> keep;
  for (SplitKScratch* p : g_splitk) {
    if (s == nullptr || p->stream == s) {
      p->buf.release();
      delete p;
    } else {
      keep.push_back(p);
    }
  }
  g_splitk.swap(keep);
}

size_t syrk_split_scratch_bytes(int n, int k) {
  (void)k;
  return sizeof(double) * (size_t)n * n + 8;
}
// occ A A^T through the split-BF16 GEMM into a dense n x n scratch, then packed.
void syrk_packed_split_bf16(hipStream_t s, int n, int k, double occ, const double* a, int lda,
                            double* p, void* scratch) {
  if (n <= 0) return;
  double* d = static_cast<double*>(scratch);
  gemm_split_bf16(s, 'N', 'T', n, n, k, occ, a, lda, a, lda, 0.0, d, n);
  pack_lower(s, n, d, n, p);
}
}  // namespace mhip
