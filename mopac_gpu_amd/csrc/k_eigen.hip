// FP64 symmetric eigensolver: parallel two-sided block Jacobi.
//
// The matrix (padded to a multiple of 64 with decoupled large diagonal
// entries) is split into 32-wide blocks.  A sweep visits every block pair once
// in round-robin order; each step
//   1. diagonalizes the 64x64 submatrix of every pair in LDS (inner cyclic
//      Jacobi, one workgroup per pair) giving an orthogonal Q_pair,
//   2. applies A <- Qt^T A Qt tile by tile (two 64^3 products per tile),
//   3. applies C <- C Qt to the eigenvector accumulator.
// In the SCF the solver starts from the previous eigenvectors (A = C^T F C is
// then nearly diagonal and one to three sweeps converge it); the first call
// starts from C = I.  Stopping test: max |A_ij| (i != j) <= tol * max|A_ii|.
#include "common.h"
#include "k_eigen.h"
#include "mfma.h"

namespace mhip {
namespace {

constexpr int BS = 32;  // block width
constexpr int SB = 64;  // pair width
constexpr int LDSW = SB + 1;

__device__ __forceinline__ int gidx_dev(int pairI, int pairJ, int local) {
  return (local < BS) ? pairI * BS + local : pairJ * BS + (local - BS);
}

__device__ __forceinline__ void rr_pair(int round, int k, int& p, int& q) {
  // circle method on 64 indices: index 0 fixed, 1..63 rotate
  auto pos = [&](int slot) -> int { return slot == 0 ? 0 : 1 + ((slot - 1 + round) % 63); };
  p = pos(k);
  q = pos(63 - k);
}

constexpr int SOLVE_THREADS = 1024;

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Rotation k of round `round` from the current submatrix (cyclic Jacobi
// angle, the smaller of the two solutions).
__device__ __forceinline__ void pair_rotation(const double (*S)[LDSW], int round, int k, int& p, int& q,
                                              double& c, double& sn) {
  rr_pair(round, k, p, q);
  if (p > q) { const int t = p; p = q; q = t; }
  const double app = S[p][p], aqq = S[q][q], apq = S[p][q];
  c = 1.0;
  sn = 0.0;
  if (fabs(apq) > 1.0e-300 && fabs(apq) > 1.0e-18 * sqrt(fabs(app * aqq))) {
    // t = tan of the rotation angle, the root of t^2 + 2 theta t - 1 = 0
    // (theta = d / h) of smaller magnitude, written without theta:
    // t = sign(d) h / (|d| + sqrt(d^2 + h^2)); one division instead of three
    const double d = aqq - app, h = 2.0 * apq;
    double t = h / (fabs(d) + sqrt(d * d + h * h));
    if (d < 0.0) t = -t;
    c = rsqrt(t * t + 1.0);
    sn = t * c;
  }
}

// One workgroup per block pair: load the 64x64 submatrix, run cyclic Jacobi
// in parallel (round-robin) order until converged or max_inner sweeps, and
// write the accumulated rotation Q.  A pair already diagonal to `skip_abs`
// is marked inactive (Q = I) and skipped by the update kernels.
//
// A round: the first wave computes the round's 32 disjoint rotations into
// LDS; then the 1024 threads each own one 2x2 block (row pair k1, column pair
// k2) of the submatrix and apply rotation k1 to its rows and k2 to its
// columns in one pass, in place (no other thread touches the block), and
// rotate two rows of Q's columns p(k2), q(k2).  Two barriers per round.
__global__ __launch_bounds__(SOLVE_THREADS) void k_jac_solve(
    const double* __restrict__ A, int lda, int nreal, const int* __restrict__ pairs,
    double* __restrict__ Qbuf, int* __restrict__ active, double tol, double skip_abs,
    int max_inner) {
  static_assert(SOLVE_THREADS == (SB / 2) * (SB / 2), "one thread per 2x2 block");
  __shared__ double S[SB][LDSW];
  __shared__ double Q[SB][LDSW];
  __shared__ double rc[32], rs[32];
  __shared__ int rp[32], rq[32];
  __shared__ double red_off[SOLVE_THREADS / 64], red_dg[SOLVE_THREADS / 64];
  const int pidx = blockIdx.x;
  const int I = pairs[2 * pidx], J = pairs[2 * pidx + 1];
  const int tid = threadIdx.x;
  const int k1 = tid >> 5, k2 = tid & 31;
  {
    // all four loads of this thread in flight before the first LDS store
    constexpr int PER = SB * SB / SOLVE_THREADS;
    double v[PER];
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int e = tid + SOLVE_THREADS * it;
      v[it] = A[(size_t)gidx_dev(I, J, e % SB) + (size_t)gidx_dev(I, J, e / SB) * lda];
    }
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int e = tid + SOLVE_THREADS * it, r = e % SB, c = e / SB;
      S[r][c] = v[it];
      Q[r][c] = (r == c) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  for (int sw = 0; sw < max_inner; ++sw) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < SB * SB; e += SOLVE_THREADS) {
      const int r = e % SB, c = e / SB;
      if (r == c) {
        if (gidx_dev(I, J, r) < nreal) dg = fmax(dg, fabs(S[r][c]));
      } else {
        off = fmax(off, fabs(S[r][c]));
      }
    }
    off = wave_max(off);
    dg = wave_max(dg);
    if ((tid & 63) == 0) {
      red_off[tid >> 6] = off;
      red_dg[tid >> 6] = dg;
    }
    __syncthreads();
    off = 0.0;
    dg = 0.0;
#pragma unroll
    for (int w = 0; w < SOLVE_THREADS / 64; ++w) {
      off = fmax(off, red_off[w]);
      dg = fmax(dg, red_dg[w]);
    }
    __syncthreads();
    if (sw == 0 && off <= skip_abs) {
      if (tid == 0) active[pidx] = 0;
      return;
    }
    if (off <= tol * fmax(dg, 1.0)) break;
    for (int round = 0; round < 63; ++round) {
      if (tid < 32) {
        int p, q;
        double c, sn;
        pair_rotation(S, round, tid, p, q, c, sn);
        rc[tid] = c; rs[tid] = sn; rp[tid] = p; rq[tid] = q;
      }
      __syncthreads();
      const int p1 = rp[k1], q1 = rq[k1], p2 = rp[k2], q2 = rq[k2];
      const double c1 = rc[k1], s1 = rs[k1], c2 = rc[k2], s2 = rs[k2];
      const double xpp = S[p1][p2], xpq = S[p1][q2], xqp = S[q1][p2], xqq = S[q1][q2];
      // rows p1, q1 (rotation k1), then columns p2, q2 (rotation k2)
      const double rpp = c1 * xpp - s1 * xqp, rpq = c1 * xpq - s1 * xqq;
      const double rqp = s1 * xpp + c1 * xqp, rqq = s1 * xpq + c1 * xqq;
      S[p1][p2] = c2 * rpp - s2 * rpq;
      S[p1][q2] = s2 * rpp + c2 * rpq;
      S[q1][p2] = c2 * rqp - s2 * rqq;
      S[q1][q2] = s2 * rqp + c2 * rqq;
      // Q <- Q J(k2) on rows 2 k1, 2 k1 + 1
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int row = 2 * k1 + h;
        const double u = Q[row][p2], v = Q[row][q2];
        Q[row][p2] = c2 * u - s2 * v;
        Q[row][q2] = s2 * u + c2 * v;
      }
      __syncthreads();
    }
  }
  double* Qo = Qbuf + (size_t)pidx * SB * SB;
  for (int e = tid; e < SB * SB; e += SOLVE_THREADS) {
    const int r = e % SB, c = e / SB;
    Qo[e] = Q[r][c];
  }
  if (tid == 0) active[pidx] = 1;
}

// Angle of the rotation that zeroes apq in [[app, apq], [apq, aqq]] (as in
// pair_rotation), with t = tan: after it the diagonal is app - t apq,
// aqq + t apq.
__device__ __forceinline__ void rotation_of(double app, double aqq, double apq, double& c, double& sn,
                                            double& t) {
  c = 1.0;
  sn = 0.0;
  t = 0.0;
  if (fabs(apq) > 1.0e-300 && fabs(apq) > 1.0e-18 * sqrt(fabs(app * aqq))) {
    const double d = aqq - app, h = 2.0 * apq;
    t = h / (fabs(d) + sqrt(d * d + h * h));
    if (d < 0.0) t = -t;
    c = rsqrt(t * t + 1.0);
    sn = t * c;
  }
}

// Thread (k1, k2) of a round: rotation k1 on rows p1, q1 and k2 on columns
// p2, q2 of the 2x2 block it owns, and on two rows of Q's columns p2, q2.
// Returns the new S[p1][q2].
__device__ __forceinline__ double round_update(double (*S)[LDSW], double (*Q)[LDSW], int k1, int p1,
                                               int q1, double c1, double s1, int p2, int q2, double c2,
                                               double s2) {
  const double xpp = S[p1][p2], xpq = S[p1][q2], xqp = S[q1][p2], xqq = S[q1][q2];
  const double rpp = c1 * xpp - s1 * xqp, rpq = c1 * xpq - s1 * xqq;
  const double rqp = s1 * xpp + c1 * xqp, rqq = s1 * xpq + c1 * xqq;
  const double npq = s2 * rpp + c2 * rpq;
  S[p1][p2] = c2 * rpp - s2 * rpq;
  S[p1][q2] = npq;
  S[q1][p2] = c2 * rqp - s2 * rqq;
  S[q1][q2] = s2 * rqp + c2 * rqq;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int row = 2 * k1 + h;
    const double u = Q[row][p2], v = Q[row][q2];
    Q[row][p2] = c2 * u - s2 * v;
    Q[row][q2] = s2 * u + c2 * v;
  }
  return npq;
}

// round_update for the cross rounds, with the elements that stay with the
// thread through them in registers: S[k1][k2] (sll) and Q's rows 2 k1,
// 2 k1 + 1 of column k2 (ql0, ql1).  Only the moving row q1 and column q2
// go through LDS; the arithmetic is round_update's, in the same order.
__device__ __forceinline__ double round_update_x(double (*S)[LDSW], double (*Q)[LDSW], int k1, int q1,
                                                 double c1, double s1, int p2, int q2, double c2, double s2,
                                                 double& sll, double& ql0, double& ql1) {
  const int p1 = k1;
  const double xpp = sll, xpq = S[p1][q2], xqp = S[q1][p2], xqq = S[q1][q2];
  const double rpp = c1 * xpp - s1 * xqp, rpq = c1 * xpq - s1 * xqq;
  const double rqp = s1 * xpp + c1 * xqp, rqq = s1 * xpq + c1 * xqq;
  const double npq = s2 * rpp + c2 * rpq;
  sll = c2 * rpp - s2 * rpq;
  S[p1][q2] = npq;
  S[q1][p2] = c2 * rqp - s2 * rqq;
  S[q1][q2] = s2 * rqp + c2 * rqq;
  {
    const double u = ql0, v = Q[2 * k1][q2];
    ql0 = c2 * u - s2 * v;
    Q[2 * k1][q2] = s2 * u + c2 * v;
  }
  {
    const double u = ql1, v = Q[2 * k1 + 1][q2];
    ql1 = c2 * u - s2 * v;
    Q[2 * k1 + 1][q2] = s2 * u + c2 * v;
  }
  return npq;
}

// The pair solve with cross rounds.  An inner sweep is, when intra[pair] is
// set, 31 rounds inside each 32-wide half (circle order on 32 indices, both
// halves at once), then 32 cross rounds: round r pairs local index k with
// 32 + (k + r) % 32, so the 1024 index pairs across the two blocks are each
// rotated once.  The sweep schedule sets intra once per sweep of the block's
// group, so every index pair of the matrix is rotated once per sweep, in
// about half the sequential rounds of a full 64-index inner sweep.
//
// Cross rounds need one barrier each: the element rotation k of round r + 1
// acts on, S[k][32 + (k + r + 1) % 32], lies in the 2x2 block of thread
// (k, k + 1) in round r, so that thread computes the next angle right after
// its own update, from the diagonals tracked in dgl (a - t apq, d + t apq
// per rotation), and publishes it in the other buffer.
__global__ __launch_bounds__(SOLVE_THREADS) void k_jac_solve_x(
    const double* __restrict__ A, int lda, int nreal, const int* __restrict__ pairs,
    const int* __restrict__ intra, double* __restrict__ Qbuf, int* __restrict__ active, double tol,
    double skip_abs, int max_inner) {
  __shared__ double S[SB][LDSW];
  __shared__ double Q[SB][LDSW];
  __shared__ double2 rot[2][32];
  __shared__ double dgl[2][SB];
  __shared__ int rp[32], rq[32];
  __shared__ double red_off[SOLVE_THREADS / 64], red_dg[SOLVE_THREADS / 64];
  const int pidx = blockIdx.x;
  const int I = pairs[2 * pidx], J = pairs[2 * pidx + 1];
  const bool do_intra = intra[pidx] != 0;
  const int tid = threadIdx.x;
  // blocks by diagonal: thread t owns (k1, k2) = (t % 32, (t % 32 + d) % 32)
  // with d = (t / 32 + 1) % 32, so the 32 blocks (k, k + 1) whose owners
  // compute the next round's angles are threads 0..31: one wave runs that
  // divergent code instead of every wave running it for two lanes
  const int k1 = tid & 31, k2 = (k1 + (((tid >> 5) + 1) & 31)) & 31;
  {
    // all four loads of this thread in flight before the first LDS store
    constexpr int PER = SB * SB / SOLVE_THREADS;
    double v[PER];
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int e = tid + SOLVE_THREADS * it;
      v[it] = A[(size_t)gidx_dev(I, J, e % SB) + (size_t)gidx_dev(I, J, e / SB) * lda];
    }
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int e = tid + SOLVE_THREADS * it, r = e % SB, c = e / SB;
      S[r][c] = v[it];
      Q[r][c] = (r == c) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  for (int sw = 0; sw < max_inner; ++sw) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < SB * SB; e += SOLVE_THREADS) {
      const int r = e % SB, c = e / SB;
      if (r == c) {
        if (gidx_dev(I, J, r) < nreal) dg = fmax(dg, fabs(S[r][c]));
      } else {
        off = fmax(off, fabs(S[r][c]));
      }
    }
    off = wave_max(off);
    dg = wave_max(dg);
    if ((tid & 63) == 0) {
      red_off[tid >> 6] = off;
      red_dg[tid >> 6] = dg;
    }
    __syncthreads();
    off = 0.0;
    dg = 0.0;
#pragma unroll
    for (int w = 0; w < SOLVE_THREADS / 64; ++w) {
      off = fmax(off, red_off[w]);
      dg = fmax(dg, red_dg[w]);
    }
    __syncthreads();
    if (sw == 0 && off <= skip_abs) {
      if (tid == 0) active[pidx] = 0;
      return;
    }
    if (off <= tol * fmax(dg, 1.0)) break;
    if (do_intra) {
      for (int round = 0; round < 31; ++round) {
        if (tid < 32) {
          // circle order on the 32 indices of half tid / 16
          const int h = tid >> 4, kk = tid & 15;
          auto pos = [&](int slot) { return slot == 0 ? 0 : 1 + ((slot - 1 + round) % 31); };
          int p = 32 * h + pos(kk), q = 32 * h + pos(31 - kk);
          if (p > q) { const int x = p; p = q; q = x; }
          double c, sn, t;
          rotation_of(S[p][p], S[q][q], S[p][q], c, sn, t);
          rot[0][tid] = make_double2(c, sn);
          rp[tid] = p;
          rq[tid] = q;
        }
        __syncthreads();
        const double2 r1 = rot[0][k1], r2 = rot[0][k2];
        (void)round_update(S, Q, k1, rp[k1], rq[k1], r1.x, r1.y, rp[k2], rq[k2], r2.x, r2.y);
        __syncthreads();
      }
    }
    // cross rounds: angles of round 0 and the diagonals after it
    if (tid < 32) {
      const int p = tid, q = 32 + tid;
      const double app = S[p][p], aqq = S[q][q], apq = S[p][q];
      double c, sn, t;
      rotation_of(app, aqq, apq, c, sn, t);
      rot[0][tid] = make_double2(c, sn);
      dgl[0][p] = app - t * apq;
      dgl[0][q] = aqq + t * apq;
    }
    __syncthreads();
    // row k1 and column k2 of this thread's block are the same in every
    // cross round: their shared S element and Q's two elements of column k2
    // stay in registers until the rounds end
    double sll = S[k1][k2], ql0 = Q[2 * k1][k2], ql1 = Q[2 * k1 + 1][k2];
    for (int r = 0; r < 32; ++r) {
      const int b = r & 1;
      const double2 r1 = rot[b][k1], r2 = rot[b][k2];
      const int q1 = 32 + ((k1 + r) & 31), q2 = 32 + ((k2 + r) & 31);
      const double npq = round_update_x(S, Q, k1, q1, r1.x, r1.y, k2, q2, r2.x, r2.y, sll, ql0, ql1);
      if (r < 31 && tid < 32) {  // thread (k1, k1 + 1)
        // rotation k1 of round r + 1 pairs k1 with q2
        const double dp = dgl[b][k1], dq = dgl[b][q2];
        double c, sn, t;
        rotation_of(dp, dq, npq, c, sn, t);
        rot[b ^ 1][k1] = make_double2(c, sn);
        dgl[b ^ 1][k1] = dp - t * npq;
        dgl[b ^ 1][q2] = dq + t * npq;
      }
      __syncthreads();
    }
    S[k1][k2] = sll;
    Q[2 * k1][k2] = ql0;
    Q[2 * k1 + 1][k2] = ql1;
    __syncthreads();
  }
  double* Qo = Qbuf + (size_t)pidx * SB * SB;
  for (int e = tid; e < SB * SB; e += SOLVE_THREADS) {
    const int r = e % SB, c = e / SB;
    Qo[e] = Q[r][c];
  }
  if (tid == 0) active[pidx] = 1;
}

__device__ __forceinline__ int gidx(int pairI, int pairJ, int local) {
  return (local < BS) ? pairI * BS + local : pairJ * BS + (local - BS);
}

// Tile t -> pair indices P1 <= P2.  The pairs [0, npo) and [npo, npairs) form
// two groups (npo == npairs: one group); tiles across the groups are not
// enumerated, since the grouped schedule never couples them.
__device__ __forceinline__ void tri_of(int t, int& P1, int& P2) {
  P2 = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((P2 + 1) * (P2 + 2) / 2 <= t) ++P2;
  while (P2 * (P2 + 1) / 2 > t) --P2;
  P1 = t - P2 * (P2 + 1) / 2;
}
__device__ __forceinline__ void tile_of(int t, int npairs, int npo, int& P1, int& P2) {
  (void)npairs;
  const int t0 = npo * (npo + 1) / 2;
  if (t < t0) {
    tri_of(t, P1, P2);
  } else {
    tri_of(t - t0, P1, P2);
    P1 += npo;
    P2 += npo;
  }
}

// A(rows of pair P1, cols of pair P2) <- Q1^T A Q2 for P1 <= P2 (the mirror
// tile is written as the transpose, keeping A exactly symmetric).
__global__ __launch_bounds__(256) void k_jac_apply_a(double* __restrict__ A, int lda, int npairs,
                                                     int npo, const int* __restrict__ pairs,
                                                     const double* __restrict__ Qbuf,
                                                     const int* __restrict__ active) {
  // blockIdx.x enumerates (P1 <= P2)
  int P1, P2;
  tile_of(blockIdx.x, npairs, npo, P1, P2);
  const bool a1 = active[P1] != 0, a2 = active[P2] != 0;
  if (!a1 && !a2) return;
  __shared__ double At[SB][LDSW];
  __shared__ double Q1[SB][LDSW];
  __shared__ double Q2[SB][LDSW];
  const int I1 = pairs[2 * P1], J1 = pairs[2 * P1 + 1];
  const int I2 = pairs[2 * P2], J2 = pairs[2 * P2 + 1];
  const int tid = threadIdx.x;
  const double* q1 = Qbuf + (size_t)P1 * SB * SB;
  const double* q2 = Qbuf + (size_t)P2 * SB * SB;
  for (int e = tid; e < SB * SB; e += 256) {
    const int r = e % SB, c = e / SB;
    At[r][c] = A[(size_t)gidx(I1, J1, r) + (size_t)gidx(I2, J2, c) * lda];
    Q1[r][c] = a1 ? q1[e] : (r == c ? 1.0 : 0.0);
    Q2[r][c] = a2 ? q2[e] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  double acc[4][4] = {};
  // T = Q1^T At : T[r][c] = sum_k Q1[k][r] At[k][c]
  for (int k = 0; k < SB; ++k) {
    double a4[4], b4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a4[i] = Q1[k][ty * 4 + i];
#pragma unroll
    for (int j = 0; j < 4; ++j) b4[j] = At[k][tx * 4 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fma(a4[i], b4[j], acc[i][j]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) At[ty * 4 + i][tx * 4 + j] = acc[i][j];
  __syncthreads();
  double out[4][4] = {};
  // Out = T Q2 : Out[r][c] = sum_k T[r][k] Q2[k][c]
  for (int k = 0; k < SB; ++k) {
    double a4[4], b4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a4[i] = At[ty * 4 + i][k];
#pragma unroll
    for (int j = 0; j < 4; ++j) b4[j] = Q2[k][tx * 4 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) out[i][j] = fma(a4[i], b4[j], out[i][j]);
  }
  if (P1 == P2) {
    // diagonal tile: symmetrize exactly
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) At[ty * 4 + i][tx * 4 + j] = out[i][j];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = ty * 4 + i, c = tx * 4 + j;
        const double v = (r == c) ? out[i][j] : 0.5 * (out[i][j] + At[c][r]);
        A[(size_t)gidx(I1, J1, r) + (size_t)gidx(I2, J2, c) * lda] = v;
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = ty * 4 + i, c = tx * 4 + j;
      const size_t gr = gidx(I1, J1, r), gc = gidx(I2, J2, c);
      A[gr + gc * lda] = out[i][j];
      A[gc + gr * lda] = out[i][j];
    }
}

// C(rows tile, cols of pair P) <- C(rows tile, cols of P) Q_P
__global__ __launch_bounds__(256) void k_jac_apply_c(double* __restrict__ C, int ldc, int nrows,
                                                     const int* __restrict__ pairs,
                                                     const double* __restrict__ Qbuf,
                                                     const int* __restrict__ active) {
  const int P = blockIdx.y;
  if (!active[P]) return;
  __shared__ double Ct[SB][LDSW];
  __shared__ double Qs[SB][LDSW];
  const int row0 = blockIdx.x * SB;
  const int I = pairs[2 * P], J = pairs[2 * P + 1];
  const double* q = Qbuf + (size_t)P * SB * SB;
  const int tid = threadIdx.x;
  for (int e = tid; e < SB * SB; e += 256) {
    const int r = e % SB, c = e / SB;
    const int gr = row0 + r;
    Ct[r][c] = (gr < nrows) ? C[(size_t)gr + (size_t)gidx(I, J, c) * ldc] : 0.0;
    Qs[r][c] = q[e];
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  double out[4][4] = {};
  for (int k = 0; k < SB; ++k) {
    double a4[4], b4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a4[i] = Ct[ty * 4 + i][k];
#pragma unroll
    for (int j = 0; j < 4; ++j) b4[j] = Qs[k][tx * 4 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) out[i][j] = fma(a4[i], b4[j], out[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gr = row0 + ty * 4 + i;
    if (gr >= nrows) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) C[(size_t)gr + (size_t)gidx(I, J, tx * 4 + j) * ldc] = out[i][j];
  }
}


// This thread's 8 elements (e = tid + 512 it) of a 64x64 rotation, all
// loads issued together; Q = I for a pair marked inactive.
constexpr int UPD_NT = 512;  // eight waves per update workgroup
__device__ __forceinline__ void stage_q(double (&v)[SB * SB / UPD_NT], const double* __restrict__ q, bool active,
                                        int tid) {
  if (active) {
#pragma unroll
    for (int it = 0; it < SB * SB / UPD_NT; ++it) v[it] = q[tid + UPD_NT * it];
  } else {
#pragma unroll
    for (int it = 0; it < SB * SB / UPD_NT; ++it) {
      const int e = tid + UPD_NT * it;
      v[it] = (e % SB == e / SB) ? 1.0 : 0.0;
    }
  }
}

// MFMA versions of the two update kernels (same data flow as above).
__global__ __launch_bounds__(UPD_NT) void k_jac_apply_a_mfma(double* __restrict__ A, int lda,
                                                          int npairs, int npo,
                                                          const int* __restrict__ pairs,
                                                          const double* __restrict__ Qbuf,
                                                          const int* __restrict__ active) {
  int P1, P2;
  tile_of(blockIdx.x, npairs, npo, P1, P2);
  const bool a1 = active[P1] != 0, a2 = active[P2] != 0;
  if (!a1 && !a2) return;
  // two 64 x 65 LDS tiles (66.5 KB) so two workgroups fit on a CU: the tile of
  // A, then Q1^T A, then the result; Q1, then Q2
  __shared__ double At[SB][LDSW];
  __shared__ double Qs[SB][LDSW];
  const int I1 = pairs[2 * P1], J1 = pairs[2 * P1 + 1];
  const int I2 = pairs[2 * P2], J2 = pairs[2 * P2 + 1];
  const int tid = threadIdx.x;
  const double* q1 = Qbuf + (size_t)P1 * SB * SB;
  const double* q2 = Qbuf + (size_t)P2 * SB * SB;
  // every global load of a stage is issued before the first LDS store, so a
  // thread waits for memory once per stage, not once per element; Q2 is
  // fetched into registers now and lands in LDS after the first product
  constexpr int PER = SB * SB / UPD_NT;  // element e = tid + 512 it: row tid % 64, column tid / 64 + 8 it
  double va[PER], vq[PER], vq2[PER];
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const int e = tid + UPD_NT * it;
    va[it] = A[(size_t)gidx(I1, J1, e % SB) + (size_t)gidx(I2, J2, e / SB) * lda];
  }
  stage_q(vq, q1, a1, tid);
  if (P1 != P2) stage_q(vq2, q2, a2, tid);
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const int e = tid + UPD_NT * it;
    At[e % SB][e / SB] = va[it];
    Qs[e % SB][e / SB] = vq[it];
  }
  __syncthreads();
  d4_t acc[2];
  for (int i = 0; i < 2; ++i) acc[i] = d4_t{0.0, 0.0, 0.0, 0.0};
  mfma_tile64_w8([&](int r, int k) { return Qs[k][r]; }, [&](int k, int c) { return At[k][c]; }, SB, acc);
  __syncthreads();
  mfma_tile64_w8_store(acc, [&](int r, int c, double v) { At[r][c] = v; });
  if (P1 != P2) {
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int e = tid + UPD_NT * it;
      Qs[e % SB][e / SB] = vq2[it];
    }
  }
  __syncthreads();
  for (int i = 0; i < 2; ++i) acc[i] = d4_t{0.0, 0.0, 0.0, 0.0};
  mfma_tile64_w8([&](int r, int k) { return At[r][k]; }, [&](int k, int c) { return Qs[k][c]; }, SB, acc);
  // stage the result in LDS and write both the tile and its mirror with
  // consecutive lanes on consecutive rows of A
  __syncthreads();
  mfma_tile64_w8_store(acc, [&](int r, int c, double v) { At[r][c] = v; });
  __syncthreads();
  if (P1 == P2) {
    for (int e = tid; e < SB * SB; e += UPD_NT) {
      const int r = e % SB, c = e / SB;
      const double v = (r == c) ? At[r][c] : 0.5 * (At[r][c] + At[c][r]);
      A[(size_t)gidx(I1, J1, r) + (size_t)gidx(I2, J2, c) * lda] = v;
    }
    return;
  }
  for (int e = tid; e < SB * SB; e += UPD_NT) {
    const int r = e % SB, c = e / SB;
    A[(size_t)gidx(I1, J1, r) + (size_t)gidx(I2, J2, c) * lda] = At[r][c];
  }
  for (int e = tid; e < SB * SB; e += UPD_NT) {
    const int c = e % SB, r = e / SB;
    A[(size_t)gidx(I2, J2, c) + (size_t)gidx(I1, J1, r) * lda] = At[r][c];
  }
}

__global__ __launch_bounds__(UPD_NT) void k_jac_apply_c_mfma(double* __restrict__ C, int ldc, int nrows,
                                                          const int* __restrict__ pairs,
                                                          const double* __restrict__ Qbuf,
                                                          const int* __restrict__ active) {
  const int P = blockIdx.y;
  if (!active[P]) return;
  __shared__ double Ct[SB][LDSW];
  __shared__ double Qs[SB][LDSW];
  const int row0 = blockIdx.x * SB;
  const int I = pairs[2 * P], J = pairs[2 * P + 1];
  const double* q = Qbuf + (size_t)P * SB * SB;
  const int tid = threadIdx.x;
  constexpr int PER = SB * SB / UPD_NT;
  double vc[PER], vq[PER];
  {
    const int gr = row0 + tid % SB;  // the same row for all of this thread's elements
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int e = tid + UPD_NT * it;
      vc[it] = (gr < nrows) ? C[(size_t)gr + (size_t)gidx(I, J, e / SB) * ldc] : 0.0;
    }
  }
  stage_q(vq, q, true, tid);
#pragma unroll
  for (int it = 0; it < PER; ++it) {
    const int e = tid + UPD_NT * it;
    Ct[e % SB][e / SB] = vc[it];
    Qs[e % SB][e / SB] = vq[it];
  }
  __syncthreads();
  d4_t acc[2];
  for (int i = 0; i < 2; ++i) acc[i] = d4_t{0.0, 0.0, 0.0, 0.0};
  mfma_tile64_w8([&](int r, int k) { return Ct[r][k]; }, [&](int k, int c) { return Qs[k][c]; }, SB, acc);
  // stage in LDS, then store with consecutive lanes on consecutive rows of C
  __syncthreads();
  mfma_tile64_w8_store(acc, [&](int r, int c, double v) { Ct[r][c] = v; });
  __syncthreads();
  for (int e = tid; e < SB * SB; e += UPD_NT) {
    const int r = e % SB, c = e / SB;
    const int gr = row0 + r;
    if (gr < nrows) C[(size_t)gr + (size_t)gidx(I, J, c) * ldc] = Ct[r][c];
  }
}

// max |A_ij| (i != j) and max |A_ii| over the leading n x n block.
// split > 0: only entries inside [0, split) x [0, split) and its complement
// count (the grouped schedule never rotates across the two groups)
__global__ __launch_bounds__(256) void k_offmax_partial(int n, const double* __restrict__ A,
                                                        int lda, int split, double* __restrict__ part) {
  const int j = blockIdx.x;
  double off = 0.0, dg = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double v = fabs(A[(size_t)i + (size_t)j * lda]);
    if (i == j) dg = v;
    else if (split <= 0 || ((i < split) == (j < split))) off = fmax(off, v);
  }
  __shared__ double r1[256], r2[256];
  r1[threadIdx.x] = off;
  r2[threadIdx.x] = dg;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + w]);
      r2[threadIdx.x] = fmax(r2[threadIdx.x], r2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * j] = r1[0];
    part[2 * j + 1] = r2[0];
  }
}

__global__ __launch_bounds__(1024) void k_offmax_final(int n, const double* __restrict__ part,
                                                       double* __restrict__ out) {
  double off = 0.0, dg = 0.0;
  for (int j = threadIdx.x; j < n; j += 1024) {
    off = fmax(off, part[2 * j]);
    dg = fmax(dg, part[2 * j + 1]);
  }
  __shared__ double r1[1024], r2[1024];
  r1[threadIdx.x] = off;
  r2[threadIdx.x] = dg;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + w]);
      r2[threadIdx.x] = fmax(r2[threadIdx.x], r2[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r1[0];
    out[1] = r2[0];
  }
}

// active pairs of the step in each group: out[0] over [0, npo), out[1] over [npo, npairs)
__global__ void k_count_active(int npairs, int npo, const int* __restrict__ active, int* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int a = 0, b = 0;
    for (int i = 0; i < npairs; ++i) {
      if (i < npo) a += active[i] != 0;
      else b += active[i] != 0;
    }
    out[0] = a;
    out[1] = b;
  }
}

// Stable ascending rank sort of the diagonal; perm[rank] = index.  One
// thread per entry across the grid; the diagonal goes through LDS in chunks.
constexpr int SORT_CHUNK = 2048;
__global__ __launch_bounds__(256) void k_sort_diag(int npad, const double* __restrict__ A, int lda,
                                                   int* __restrict__ perm, double* __restrict__ eig) {
  __shared__ double d[SORT_CHUNK];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const double di = (i < npad) ? A[(size_t)i + (size_t)i * lda] : 0.0;
  int rank = 0;
  for (int j0 = 0; j0 < npad; j0 += SORT_CHUNK) {
    const int nj = min(SORT_CHUNK, npad - j0);
    for (int j = threadIdx.x; j < nj; j += 256) d[j] = A[(size_t)(j0 + j) + (size_t)(j0 + j) * lda];
    __syncthreads();
    for (int j = 0; j < nj; ++j) {
      const double dj = d[j];
      rank += (dj < di) || (dj == di && j0 + j < i);
    }
    __syncthreads();
  }
  if (i < npad) {
    perm[rank] = i;
    eig[rank] = di;
  }
}

__global__ void k_permute_cols(int nrows, int ncols, const double* __restrict__ src, int lds,
                               const int* __restrict__ perm, double* __restrict__ dst, int ldd) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (r >= nrows || k >= ncols) return;
  dst[(size_t)r + (size_t)k * ldd] = src[(size_t)r + (size_t)perm[k] * lds];
}

__global__ void k_set_identity(int npad, double* __restrict__ C, int ldc) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= npad) return;
  C[(size_t)r + (size_t)c * ldc] = (r == c) ? 1.0 : 0.0;
}

// Largest-|c| coefficient of each column made positive (first occurrence wins).
__global__ __launch_bounds__(256) void k_phase_lock(int n, double* __restrict__ C, int ldc) {
  const int col = blockIdx.x;
  double* v = C + (size_t)col * ldc;
  double best = 0.0, sgn = 0.0;
  int bi = n;
  for (int r = threadIdx.x; r < n; r += 256) {
    const double a = fabs(v[r]);
    if (a > best) { best = a; sgn = v[r]; bi = r; }
  }
  __shared__ double sb[256], ss[256];
  __shared__ int si[256];
  sb[threadIdx.x] = best; ss[threadIdx.x] = sgn; si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double b2 = sb[threadIdx.x + w];
      const int i2 = si[threadIdx.x + w];
      if (b2 > sb[threadIdx.x] || (b2 == sb[threadIdx.x] && i2 < si[threadIdx.x])) {
        sb[threadIdx.x] = b2; ss[threadIdx.x] = ss[threadIdx.x + w]; si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  const bool flip = ss[0] < 0.0;
  if (flip)
    for (int r = threadIdx.x; r < n; r += 256) v[r] = -v[r];
}

}  // namespace

int g_jacobi_schedule = 1;

void JacobiPlan::build(int n_in, int split) {
  // buffers below may be re-allocated: no C update of an earlier solve may
  // still be reading them
  if (cstream) MHIP_CHECK(hipStreamSynchronize(cstream));
  n = n_in;
  npad = ((n + SB - 1) / SB) * SB;
  nb = npad / BS;
  npairs = nb / 2;
  // circle method over the blocks [b0, b0 + m): round r, pair k
  auto circle = [](int b0, int m, int r, int k, int& a, int& b) {
    auto pos = [&](int slot) { return slot == 0 ? 0 : 1 + ((slot - 1 + r) % (m - 1)); };
    a = b0 + pos(k);
    b = b0 + pos(m - 1 - k);
    if (a > b) std::swap(a, b);
  };
  const int nbo = (split > 0 && split % BS == 0) ? split / BS : 0;
  const int nbv = nb - nbo;
  grouped = nbo >= 2 && nbv >= 2 && nbo % 2 == 0 && nbv % 2 == 0;
  this->split = grouped ? split : 0;
  npo = grouped ? nbo / 2 : npairs;
  std::vector<int> h;
  if (grouped) {
    // two independent problems (blocks [0, nbo) and [nbo, nb)) in every
    // step; the smaller one repeats its rounds while the larger finishes
    nsteps = std::max(nbo, nbv) - 1;
    h.resize((size_t)nsteps * npairs * 2);
    for (int step = 0; step < nsteps; ++step) {
      int k = 0;
      for (int q = 0; q < nbo / 2; ++q, ++k)
        circle(0, nbo, step % (nbo - 1), q, h[((size_t)step * npairs + k) * 2],
               h[((size_t)step * npairs + k) * 2 + 1]);
      for (int q = 0; q < nbv / 2; ++q, ++k)
        circle(nbo, nbv, step % (nbv - 1), q, h[((size_t)step * npairs + k) * 2],
               h[((size_t)step * npairs + k) * 2 + 1]);
    }
  } else {
    nsteps = nb - 1;
    h.resize((size_t)nsteps * npairs * 2);
    for (int step = 0; step < nsteps; ++step)
      for (int k = 0; k < npairs; ++k)
        circle(0, nb, step, k, h[((size_t)step * npairs + k) * 2], h[((size_t)step * npairs + k) * 2 + 1]);
  }
  pairs.ensure(h.size() * sizeof(int));
  MHIP_CHECK(hipMemcpy(pairs.ptr, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
  // intra-block rounds at the first step of each sweep of a pair's group
  std::vector<int> fl((size_t)nsteps * npairs, 0);
  for (int step = 0; step < nsteps; ++step)
    for (int k = 0; k < npairs; ++k) {
      const bool first = grouped ? (k < npo ? step % (nbo - 1) == 0 : step % (nbv - 1) == 0) : step == 0;
      fl[(size_t)step * npairs + k] = first ? 1 : 0;
    }
  intra_host = fl;
  intra.ensure(fl.size() * sizeof(int));
  MHIP_CHECK(hipMemcpy(intra.ptr, fl.data(), fl.size() * sizeof(int), hipMemcpyHostToDevice));
  qbuf.ensure((size_t)nsteps * npairs * SB * SB * sizeof(double));
  part.ensure((size_t)2 * npad * sizeof(double) + 64);
  perm.ensure((size_t)npad * sizeof(int));
  active.ensure((size_t)nsteps * npairs * sizeof(int));
  if (!cstream) MHIP_CHECK(hipStreamCreateWithFlags(&cstream, hipStreamNonBlocking));
  if (!ev_join) MHIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  while ((int)ev_solved.size() < nsteps) {
    hipEvent_t a, b;
    MHIP_CHECK(hipEventCreateWithFlags(&a, hipEventDisableTiming));
    MHIP_CHECK(hipEventCreateWithFlags(&b, hipEventDisableTiming));
    ev_solved.push_back(a);
    ev_cdone.push_back(b);
  }
  cdone_pending.assign(ev_cdone.size(), 0);
  counts.ensure((size_t)2 * nb * sizeof(int));
}

void jacobi_set_identity(hipStream_t s, int npad, double* C, int ldc) {
  hipLaunchKernelGGL(k_set_identity, dim3((npad + 255) / 256, npad), dim3(256), 0, s, npad, C, ldc);
  MHIP_LAUNCH_CHECK();
}

int jacobi_sweeps(hipStream_t s, JacobiPlan& plan, double* A, double* C, double tol, int max_sweeps,
                  double* host_scal, double* dev_scal, int* sweeps_done) {
  const int lda = plan.npad;
  int done = 0;
  hipStream_t cs = plan.cstream;
  // the C updates start after whatever the caller queued on C
  MHIP_CHECK(hipEventRecord(plan.ev_join, s));
  MHIP_CHECK(hipStreamWaitEvent(cs, plan.ev_join, 0));
  // and the caller's later work on C waits for the last of them
  auto join = [&] {
    MHIP_CHECK(hipEventRecord(plan.ev_join, cs));
    MHIP_CHECK(hipStreamWaitEvent(s, plan.ev_join, 0));
  };
  for (int sw = 0; sw <= max_sweeps; ++sw) {
    hipLaunchKernelGGL(k_offmax_partial, dim3(plan.n), dim3(256), 0, s, plan.n, A, lda,
                       plan.grouped ? plan.split : 0, plan.part.as<double>());
    MHIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_offmax_final, dim3(1), dim3(1024), 0, s, plan.n, plan.part.as<double>(),
                       dev_scal);
    MHIP_LAUNCH_CHECK();
    MHIP_CHECK(hipMemcpyAsync(host_scal, dev_scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipStreamSynchronize(s));
    const double thr = tol * std::max(1.0, host_scal[1]);
    if (host_scal[0] <= thr) break;
    // pairs whose whole 64x64 block is already this diagonal are left alone
    const double skip_abs = 0.1 * thr;
    if (sw == max_sweeps) {
      join();
      if (sweeps_done) *sweeps_done = done;
      return 1;
    }
    const int nsteps = plan.nsteps;
    if (plan.timing && plan.ev.size() < (size_t)(6 * nsteps)) {
      for (auto e : plan.ev) (void)hipEventDestroy(e);
      plan.ev.assign(6 * nsteps, nullptr);
      for (auto& e : plan.ev) MHIP_CHECK(hipEventCreate(&e));
    }
    const int npv = plan.npairs - plan.npo;
    const int ntiles = plan.npo * (plan.npo + 1) / 2 + npv * (npv + 1) / 2;
    for (int step = 0; step < nsteps; ++step) {
      const int* pr = plan.pairs.as<int>() + (size_t)step * plan.npairs * 2;
      double* q = plan.qbuf.as<double>() + (size_t)step * plan.npairs * SB * SB;
      int* act = plan.active.as<int>() + (size_t)step * plan.npairs;
      // the slot's previous C update (one sweep earlier) must have read it
      if (plan.cdone_pending[step]) MHIP_CHECK(hipStreamWaitEvent(s, plan.ev_cdone[step], 0));
      if (plan.timing) MHIP_CHECK(hipEventRecord(plan.ev[6 * step + 0], s));
      if (g_jacobi_schedule == 1) {
        const int* fl = plan.intra.as<int>() + (size_t)step * plan.npairs;
        hipLaunchKernelGGL(k_jac_solve_x, dim3(plan.npairs), dim3(SOLVE_THREADS), 0, s, A, lda, plan.n, pr, fl,
                           q, act, 1.0e-14, skip_abs, plan.max_inner);
      } else {
        hipLaunchKernelGGL(k_jac_solve, dim3(plan.npairs), dim3(SOLVE_THREADS), 0, s, A, lda, plan.n, pr, q,
                           act, 1.0e-14, skip_abs, plan.max_inner);
      }
      MHIP_LAUNCH_CHECK();
      if (plan.timing) {
        MHIP_CHECK(hipEventRecord(plan.ev[6 * step + 1], s));
        hipLaunchKernelGGL(k_count_active, dim3(1), dim3(64), 0, s, plan.npairs, plan.npo, act,
                           plan.counts.as<int>() + 2 * step);
      }
      if (g_kernel_variant == 1) {
        if (plan.timing) MHIP_CHECK(hipEventRecord(plan.ev[6 * step + 2], s));
        hipLaunchKernelGGL(k_jac_apply_a_mfma, dim3(ntiles), dim3(UPD_NT), 0, s, A, lda, plan.npairs, plan.npo, pr, q,
                           act);
        MHIP_LAUNCH_CHECK();
        if (plan.timing) MHIP_CHECK(hipEventRecord(plan.ev[6 * step + 3], s));
        // C <- C Q on the side stream once this step's A update is done, so
        // it overlaps the next step's pair solves, which leave most CUs idle
        MHIP_CHECK(hipEventRecord(plan.ev_solved[step], s));
        MHIP_CHECK(hipStreamWaitEvent(cs, plan.ev_solved[step], 0));
        if (plan.timing) MHIP_CHECK(hipEventRecord(plan.ev[6 * step + 4], cs));
        hipLaunchKernelGGL(k_jac_apply_c_mfma, dim3(plan.npad / SB, plan.npairs), dim3(UPD_NT), 0, cs, C, lda,
                           plan.npad, pr, q, act);
        MHIP_LAUNCH_CHECK();
        if (plan.timing) MHIP_CHECK(hipEventRecord(plan.ev[6 * step + 5], cs));
      } else {
        hipLaunchKernelGGL(k_jac_apply_a, dim3(ntiles), dim3(256), 0, s, A, lda, plan.npairs, plan.npo, pr, q,
                           act);
        MHIP_LAUNCH_CHECK();
        MHIP_CHECK(hipEventRecord(plan.ev_solved[step], s));
        MHIP_CHECK(hipStreamWaitEvent(cs, plan.ev_solved[step], 0));
        hipLaunchKernelGGL(k_jac_apply_c, dim3(plan.npad / SB, plan.npairs), dim3(256), 0, cs, C, lda,
                           plan.npad, pr, q, act);
        MHIP_LAUNCH_CHECK();
      }
      MHIP_CHECK(hipEventRecord(plan.ev_cdone[step], cs));
      plan.cdone_pending[step] = 1;
    }
    if (plan.timing && g_kernel_variant == 1) {
      MHIP_CHECK(hipStreamSynchronize(cs));
      MHIP_CHECK(hipStreamSynchronize(s));
      std::vector<int> act(2 * nsteps);
      MHIP_CHECK(hipMemcpy(act.data(), plan.counts.ptr, sizeof(int) * 2 * nsteps, hipMemcpyDeviceToHost));
      for (int step = 0; step < nsteps; ++step) {
        // tiles processed: those of each group with at least one active pair
        const long long u0 = plan.npo - act[2 * step], u1 = npv - act[2 * step + 1];
        plan.active_pairs += act[2 * step] + act[2 * step + 1];
        // a full inner sweep is 2016 rotations; cross rounds are 32 x 32,
        // plus 2 x 496 inside the blocks where the step's intra flag is set
        for (int g = 0; g < 2; ++g) {
          const int first = g == 0 ? 0 : plan.npo;
          if (first >= plan.npairs) continue;
          const long long per =
              g_jacobi_schedule == 1
                  ? 1024 + (plan.intra_host[(size_t)step * plan.npairs + first] ? 992 : 0)
                  : 2016LL * plan.max_inner;
          plan.active_rotations += per * act[2 * step + g];
        }
        plan.active_tiles += (long long)ntiles - u0 * (u0 + 1) / 2 - u1 * (u1 + 1) / 2;
        float a = 0.f, b = 0.f, c2 = 0.f;
        MHIP_CHECK(hipEventElapsedTime(&a, plan.ev[6 * step + 0], plan.ev[6 * step + 1]));
        MHIP_CHECK(hipEventElapsedTime(&b, plan.ev[6 * step + 2], plan.ev[6 * step + 3]));
        MHIP_CHECK(hipEventElapsedTime(&c2, plan.ev[6 * step + 4], plan.ev[6 * step + 5]));
        plan.t_solve_ms += a;
        plan.t_apply_ms += b;
        plan.t_applyc_ms += c2;
        plan.n_solve++;
        plan.n_apply++;
        plan.n_applyc++;
      }
    }
    ++done;
  }
  join();
  if (sweeps_done) *sweeps_done = done;
  return 0;
}

void jacobi_sort(hipStream_t s, JacobiPlan& plan, const double* A, const double* C, double* Cout,
                 double* eig_dev) {
  const int npad = plan.npad;
  hipLaunchKernelGGL(k_sort_diag, dim3((npad + 255) / 256), dim3(256), 0, s, npad, A, npad, plan.perm.as<int>(),
                     eig_dev);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_permute_cols, dim3((npad + 255) / 256, npad), dim3(256), 0, s, npad, npad, C,
                     npad, plan.perm.as<int>(), Cout, npad);
  MHIP_LAUNCH_CHECK();
}

void phase_lock(hipStream_t s, int n, double* C, int ldc) {
  hipLaunchKernelGGL(k_phase_lock, dim3(n), dim3(256), 0, s, n, C, ldc);
  MHIP_LAUNCH_CHECK();
}

}  // namespace mhip
