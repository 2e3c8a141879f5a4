// Batched SCF: thousands of small molecules of uneven size in one captured
// hipGraph.  Every kernel covers the whole batch (a workgroup per molecule,
// or per (molecule, atom / row)); molecules that have converged drop out by a
// flag, so the graph is the same for every replay.  The control logic of
// iter (src/SCF/iter.F90:340-1041: level shift, convergence test, recompute
// with shift 0, ready/iredy) runs on the device, one thread per molecule, so
// no scalar returns to the host until the whole batch is done.
//
// Per iteration: control (shift, mode) -> fock2 (batched) -> helect ->
// control (energy, test) -> eigensolve + density (one workgroup per
// molecule: warm-started cyclic Jacobi in parallel order, the packed matrix
// in LDS, and the eigenvectors too where they fit) -> cnvg (batched) -> control (PL, frontier levels, newdg) ->
// Pa = P / 2.  The eigensolve's workgroup size per size class comes from a
// measured table (mhip_batch_tune).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.h"
#include "k_fock.h"

namespace mhip {
namespace {

struct BatchMol {
  int n, nocc, numat;
  double* f;
  double* p;
  double* pa;
  double* pold;
  double* p1;
  double* C;
  double* A;
  double* T;
  double* eig;
};

struct BatchState {
  int niter, iredy, ready, done, converged, newdg, oknewd, recompute, have_c, do_eig, mode, itrmax;
  double ten, shift, shfmax, shfto, shftbo, shiftb, pl, eold, diff, sellim, escf, ee;
  double eig_homo, eig_lumo, enuclr, atheat, selcon, pltest;
  // eigensolver work of this molecule since the batch was reset: applied
  // rotations, and FP64 flops of the method (12 n per rotation: 6 n on the
  // symmetric A, 6 n on C; warm start 2 n^3 + 2 n tri(n); density
  // 2 tri(n) nocc)
  long long eig_rot;
  double eig_flops;
};

__device__ __forceinline__ double beig_flops(int n, int nocc, long long nrot, bool warm) {
  const double dn = n, tn = 0.5 * dn * (dn + 1.0);
  return 12.0 * dn * (double)nrot + (warm ? 2.0 * dn * dn * dn + 2.0 * dn * tn : 0.0) + 2.0 * tn * nocc;
}

constexpr double kEnrgy = 23.060547830619029;  // fpc_9, eV -> kcal/mol

// ---- control, part 1: shift and Fock mode (iter.F90:340-420) ---------------
__global__ void k_bctl_pre(int nmol, BatchState* __restrict__ st, int* __restrict__ act,
                           int* __restrict__ mode, int* __restrict__ niter_out,
                           double* __restrict__ shift_out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  BatchState s = st[m];
  act[m] = !s.done;
  if (s.done) return;
  if (s.recompute) {
    // label 320 again: F from H with the shift set to zero, same iteration
    s.mode = kInitH;
  } else {
    s.niter += 1;
    const double bshift = -80.0;
    s.mode = kInitH;
    if (fabs(s.shift) > 1e-10) {
      if (s.niter > 1) {
        if (s.newdg) {
          if (s.diff > 0.0) {
            s.shift = 1.0;
            if (s.diff > 1) s.newdg = 0;
          } else {
            s.shift = -0.1;
          }
        } else {
          s.shift = s.ten + s.eig_lumo - s.eig_homo + s.shift;
        }
        if (s.diff > 0.0) {
          if (s.shift > 4.0) s.shfmax = 4.5;
          if (s.shift > s.shfmax) s.shfmax = fmax(s.shfmax - 0.5, 0.0);
        }
        s.shift = fmax(-20.0, fmin(s.shfmax, s.shift));
        if (fabs(s.shift - s.shfmax) < 1e-5) s.shfmax += 0.01;
        (void)bshift;
      }
      if (s.pl > 0.005) {
        s.shftbo = s.shiftb;
        s.shfto = s.shift;
      } else {
        s.shiftb = s.shftbo;
        s.shift = s.shfto;
      }
      s.mode = kInitShift;
    }
  }
  mode[m] = s.mode;
  niter_out[m] = s.niter;
  shift_out[m] = s.shift;
  st[m] = s;
}

// ---- control, part 2: energy and convergence test (iter.F90:640-700) -------
__global__ void k_bctl_mid(int nmol, BatchState* __restrict__ st, const double* __restrict__ ee,
                           int* __restrict__ do_eig) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  BatchState s = st[m];
  do_eig[m] = 0;
  if (s.done) return;
  s.ee = ee[m];
  if (!isfinite(s.ee)) {
    s.done = 1;
    s.converged = 0;
    st[m] = s;
    return;
  }
  s.escf = (s.ee + s.enuclr) * kEnrgy + s.atheat;
  if (s.niter >= s.itrmax) {
    s.done = 1;
    st[m] = s;
    return;
  }
  const bool was_recompute = s.recompute != 0;
  s.recompute = 0;
  s.diff = s.escf - s.eold;
  if (s.diff > 0) s.ten = s.ten - 1.0;
  else s.ten = s.ten * 0.975 + 0.05;
  s.sellim = fmax(s.selcon, 1e-15 * fmax(fabs(s.ee), 1.0));
  if (s.niter > 4 && (s.pl == 0.0 || (s.pl < s.pltest && fabs(s.diff) < s.sellim)) && s.ready) {
    if (fabs(s.shift) < 1e-10) {
      s.converged = 1;
      s.done = 1;
      st[m] = s;
      return;
    }
    s.shift = 0.0;
    s.shiftb = 0.0;
    s.recompute = 1;
    st[m] = s;
    return;
  }
  (void)was_recompute;
  s.ready = s.iredy > 0 && (fabs(s.diff) < s.sellim * 10.0 || s.pl == 0.0);
  s.iredy += 1;
  s.eold = s.escf;
  do_eig[m] = 1;
  st[m] = s;
}

// ---- eigensolve + density, one workgroup per molecule ------------------------
__device__ __forceinline__ int rr_index(int idx, int r, int np) {
  return idx == 0 ? 0 : 1 + ((r + idx - 1) % (np - 1));
}

template <int NT>
__device__ double wg_max(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// maxima of two values over the workgroup: within each wave by lane
// exchange, then across the NT / 64 waves through red (2 NT / 64 doubles)
template <int NT>
__device__ void wg_max2(double& a, double& b, double* red) {
  for (int o = 32; o > 0; o >>= 1) {
    a = fmax(a, __shfl_xor(a, o));
    b = fmax(b, __shfl_xor(b, o));
  }
  constexpr int NW = NT / 64;
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[NW + w] = b;
  }
  __syncthreads();
  a = red[0];
  b = red[NW];
  for (int i = 1; i < NW; ++i) {
    a = fmax(a, red[i]);
    b = fmax(b, red[NW + i]);
  }
  __syncthreads();
}

constexpr int kMaxBatchN = 256;

template <int NT>
__global__ __launch_bounds__(NT) void k_beig(const BatchMol* __restrict__ mols, const int* __restrict__ do_eig,
                                             BatchState* __restrict__ st, const int* __restrict__ list,
                                             double tol) {
  const int m = list[blockIdx.x];
  if (!do_eig[m]) return;
  const BatchMol& M = mols[m];
  const int n = M.n, tid = threadIdx.x;
  __shared__ double red[NT];
  __shared__ double cs[kMaxBatchN / 2], sn[kMaxBatchN / 2];
  __shared__ int pp[kMaxBatchN / 2], qq[kMaxBatchN / 2];
  __shared__ int rank[kMaxBatchN];
  __shared__ int nrot;
  if (tid == 0) nrot = 0;
  double* A = M.A;
  double* C = M.C;
  double* T = M.T;
  const bool warm = st[m].have_c != 0;
  // eigenvectors_LAPACK: F(i,i) += i * 1e-10, in place
  for (int i = tid; i < n; i += NT) M.f[(long long)i * (i + 1) / 2 + i] += (double)(i + 1) * 1.0e-10;
  __syncthreads();
  auto F = [&](int r, int c) -> double {
    return r >= c ? M.f[(long long)r * (r + 1) / 2 + c] : M.f[(long long)c * (c + 1) / 2 + r];
  };
  if (warm) {
    // T = F C, A = C^T T
    for (int e = tid; e < n * n; e += NT) {
      const int r = e % n, c = e / n;
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += F(r, k) * C[k + c * n];
      T[e] = s;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += NT) {
      const int r = e % n, c = e / n;
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += C[k + r * n] * T[k + c * n];
      A[e] = s;
    }
  } else {
    for (int e = tid; e < n * n; e += NT) {
      const int r = e % n, c = e / n;
      A[e] = F(r, c);
      C[e] = (r == c) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  const int np = (n + 1) & ~1;
  const int npair = np / 2;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < n * n; e += NT) {
      const int r = e % n, c = e / n;
      if (r == c) dg = fmax(dg, fabs(A[e]));
      else off = fmax(off, fabs(A[e]));
    }
    off = wg_max<NT>(off, red);
    dg = wg_max<NT>(dg, red);
    if (off <= tol * fmax(dg, 1.0)) break;
    for (int r = 0; r < np - 1; ++r) {
      for (int k = tid; k < npair; k += NT) {
        int p = rr_index(k, r, np), q = rr_index(np - 1 - k, r, np);
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double app = A[p + p * n], aqq = A[q + q * n], apq = A[p + q * n];
          if (fabs(apq) > 1.0e-300 && fabs(apq) > 1.0e-18 * sqrt(fabs(app * aqq))) {
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
            atomicAdd(&nrot, 1);
          }
        }
        pp[k] = p;
        qq[k] = q;
        cs[k] = c;
        sn[k] = s;
      }
      __syncthreads();
      // rows p, q of A
      for (int e = tid; e < npair * n; e += NT) {
        const int k = e / n, col = e % n;
        const int p = pp[k], q = qq[k];
        if (q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double x = A[p + col * n], y = A[q + col * n];
        A[p + col * n] = c * x - s * y;
        A[q + col * n] = s * x + c * y;
      }
      __syncthreads();
      // columns p, q of A and of C
      for (int e = tid; e < npair * n; e += NT) {
        const int k = e / n, row = e % n;
        const int p = pp[k], q = qq[k];
        if (q >= n || sn[k] == 0.0) continue;
        const double c = cs[k], s = sn[k];
        const double x = A[row + p * n], y = A[row + q * n];
        A[row + p * n] = c * x - s * y;
        A[row + q * n] = s * x + c * y;
        const double u = C[row + p * n], v = C[row + q * n];
        C[row + p * n] = c * u - s * v;
        C[row + q * n] = s * u + c * v;
      }
      __syncthreads();
    }
  }
  // ascending order, ties by index
  for (int i = tid; i < n; i += NT) {
    const double ai = A[i + i * n];
    int rk = 0;
    for (int j = 0; j < n; ++j) {
      const double aj = A[j + j * n];
      rk += (aj < ai || (aj == ai && j < i)) ? 1 : 0;
    }
    rank[i] = rk;
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += NT) {
    const int r = e % n, c = e / n;
    T[r + rank[c] * n] = C[e];
  }
  for (int i = tid; i < n; i += NT) M.eig[rank[i]] = A[i + i * n];
  __syncthreads();
  for (int e = tid; e < n * n; e += NT) C[e] = T[e];
  __syncthreads();
  // P = 2 C_occ C_occ^T, packed
  const long long mp = (long long)n * (n + 1) / 2;
  for (long long q = tid; q < mp; q += NT) {
    int r = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
    while ((long long)(r + 1) * (r + 2) / 2 <= q) ++r;
    while ((long long)r * (r + 1) / 2 > q) --r;
    const int c = (int)(q - (long long)r * (r + 1) / 2);
    double s = 0.0;
    for (int k = 0; k < M.nocc; ++k) s += C[r + k * n] * C[c + k * n];
    M.p[q] = 2.0 * s;
  }
  if (tid == 0) {
    st[m].eig_rot += nrot;
    st[m].eig_flops += beig_flops(n, M.nocc, nrot, warm);
    st[m].have_c = 1;
  }
}

// LDS form of the batched eigensolver: the same warm-started cyclic Jacobi in
// parallel (round-robin) order, with the working matrix A = C^T F C kept as
// its lower triangle, packed, in LDS (np (np+1) / 2 doubles, np = n rounded
// up to even; row n of an odd n is a zero row that no rotation touches), and
// C in LDS too when both fit (CLDS).  One round is two phases: the angles of
// its np / 2 disjoint pairs, then every 2x2 block (pair k1, pair k2) of A
// turned by both rotations together (rows by k1, columns by k2), and, with
// CLDS, the columns of C.  Without CLDS the rounds log their (c, s) in global
// memory and C is turned once per sweep: A is parked in global memory, C
// passes through the freed LDS in chunks of rows, each chunk taking the
// sweep's rounds in order (the same operations on every element as turning
// C round by round), then A comes back.
// A rotation whose |a_pq| is under 1e-2 tol max(|diag|, 1) is skipped: the
// sweep loop stops once every off-diagonal element is under tol max(|diag|,
// 1), so such an element already meets it.  The warm start stages F in the
// LDS that A later takes: T' = (F C)^T in global memory (lanes along rows of
// F, C read as wave-uniform columns), then A = C^T T' with lanes along the
// columns of T' (contiguous) and four rows per thread.
__device__ __forceinline__ int tri(int i) { return (i * (i + 1)) >> 1; }
// rr_index for idx, r < np - 1 (the sum wraps at most once)
__device__ __forceinline__ int rr_pos(int idx, int r, int np) {
  if (idx == 0) return 0;
  const int t = r + idx - 1;
  return 1 + (t >= np - 1 ? t - (np - 1) : t);
}
__device__ __forceinline__ int sym_at(int a, int b) { return a >= b ? tri(a) + b : tri(b) + a; }
// row of packed lower-triangle index e (exact for e < 2^22)
__device__ __forceinline__ int tri_row(int e) {
  int i = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
  if (tri(i + 1) <= e) ++i;
  if (tri(i) > e) --i;
  return i;
}

// C <- C R^T for the logged rounds of one sweep, in chunks of at most NT rows
// staged in LDS (ch, room doubles, column-major per chunk); thread (row,
// slot) turns its row for the pairs slot, slot + NT / rows, ... of every
// round, while the pairs of the next round are staged.
template <int NT>
__device__ void beig_flush_c(double* __restrict__ C, const double* __restrict__ lg, int n, int np, double* ch,
                             int room, double2 (*csn2)[kMaxBatchN / 2], int2 (*off2)[kMaxBatchN / 2]) {
  const int tid = threadIdx.x, npair = np / 2, nround = np - 1;
  const int R = min(min(n, room / n), NT);
  for (int r0 = 0; r0 < n; r0 += R) {
    const int rows = min(R, n - r0);
    // round r's (c, s) for the pairs k = tid + j NT are fetched into
    // registers two rounds ahead and put into LDS one round ahead
    constexpr int KP = (kMaxBatchN / 2 + NT - 1) / NT;
    double fc[KP], fs[KP];
    auto fetch = [&](int r) {
#pragma unroll
      for (int j = 0; j < KP; ++j) {
        const int k = tid + j * NT;
        if (k < npair) {
          fc[j] = lg[2 * (r * npair + k)];
          fs[j] = lg[2 * (r * npair + k) + 1];
        }
      }
    };
    auto put = [&](int r) {
      const int b = r & 1;
#pragma unroll
      for (int j = 0; j < KP; ++j) {
        const int k = tid + j * NT;
        if (k < npair) {
          int p = rr_pos(k, r, np), q = rr_pos(np - 1 - k, r, np);
          if (p > q) { const int t = p; p = q; q = t; }
          csn2[b][k] = make_double2(fc[j], q < n ? fs[j] : 0.0);
          off2[b][k] = make_int2(p * rows, q * rows);
        }
      }
    };
    for (int e = tid; e < rows * n; e += NT) {
      const int row = e % rows, col = e / rows;
      ch[row + col * rows] = C[r0 + row + (size_t)col * n];
    }
    fetch(0);
    put(0);
    if (nround > 1) fetch(1);
    __syncthreads();
    const int kstep = NT / rows, row = tid % rows, slot = tid / rows;
    for (int r = 0; r < nround; ++r) {
      if (r + 1 < nround) put(r + 1);
      if (r + 2 < nround) fetch(r + 2);
      const int b = r & 1;
      if (slot < kstep) {
        for (int k = slot; k < npair; k += kstep) {
          const double2 a = csn2[b][k];
          if (a.y != 0.0) {
            const int2 o = off2[b][k];
            double* cp = ch + o.x + row;
            double* cq = ch + o.y + row;
            const double u = *cp, v = *cq;
            *cp = a.x * u - a.y * v;
            *cq = a.y * u + a.x * v;
          }
        }
      }
      __syncthreads();
    }
    for (int e = tid; e < rows * n; e += NT) {
      const int row2 = e % rows, col = e / rows;
      C[r0 + row2 + (size_t)col * n] = ch[row2 + col * rows];
    }
    __syncthreads();
  }
}

template <int NT, bool CLDS>
__global__ __launch_bounds__(NT) void k_beig_lds(const BatchMol* __restrict__ mols, const int* __restrict__ do_eig,
                                                 BatchState* __restrict__ st, const int* __restrict__ list,
                                                 double tol, int room) {
  const int m = list[blockIdx.x];
  if (!do_eig[m]) return;
  const BatchMol& M = mols[m];
  const int n = M.n, tid = threadIdx.x;
  const int np = (n + 1) & ~1;
  const int npair = np / 2;
  __shared__ double red[2 * (NT / 64)];
  __shared__ double2 csn2[2][kMaxBatchN / 2];  // (c, s) of every pair; two buffers for the flush
  __shared__ int4 pq4[kMaxBatchN / 2];         // p, q, tri(p), tri(q)
  __shared__ int2 off2[2][kMaxBatchN / 2];     // the flush's column offsets
  __shared__ int rank[kMaxBatchN];
  __shared__ int rotated, nrot;
  if (tid == 0) nrot = 0;
  double2* csn = csn2[0];
  extern __shared__ double dyn[];
  double* A = dyn;                               // tri(np) packed lower triangle
  double* C = CLDS ? dyn + tri(np) : M.C;        // n x n, column-major
  double* Tt = M.T;                              // n x n scratch, global
  const bool warm = st[m].have_c != 0;
  const int na = tri(np);
  // without CLDS: the sweep's (c, s) log, np (np-1) doubles, then A parked
  double* lg = M.A;
  double* park = M.A + (size_t)np * (np - 1);
  // eigenvectors_LAPACK: F(i,i) += i * 1e-10, in place
  for (int i = tid; i < n; i += NT) M.f[tri(i) + i] += (double)(i + 1) * 1.0e-10;
  __syncthreads();
  // F (packed, same layout as A's first n rows) into LDS, zero row n
  for (int e = tid; e < na; e += NT) A[e] = e < tri(n) ? M.f[e] : 0.0;
  if (CLDS && warm)
    for (int e = tid; e < n * n; e += NT) C[e] = M.C[e];
  __syncthreads();
  if (warm) {
    // T'(c, r) = sum_k F(r, k) C(k, c): four columns c per thread (the
    // columns past n repeat column n - 1 and are not stored)
    const int ncb = (n + 3) >> 2;
    for (int w = tid; w < n * ncb; w += NT) {
      const int r = w % n, c0 = (w / n) * 4;
      const int rb = tri(r);
      const double* q0 = C + (size_t)c0 * n;
      const double* q1 = C + (size_t)min(c0 + 1, n - 1) * n;
      const double* q2 = C + (size_t)min(c0 + 2, n - 1) * n;
      const double* q3 = C + (size_t)min(c0 + 3, n - 1) * n;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll 4
      for (int k = 0; k < n; ++k) {
        const double f = A[k <= r ? rb + k : tri(k) + r];
        s0 += f * q0[k];
        s1 += f * q1[k];
        s2 += f * q2[k];
        s3 += f * q3[k];
      }
      double* t = Tt + c0 + (size_t)r * n;
      t[0] = s0;
      if (c0 + 1 < n) t[1] = s1;
      if (c0 + 2 < n) t[2] = s2;
      if (c0 + 3 < n) t[3] = s3;
    }
    __syncthreads();
    // A(i, j) = sum_k C(k, i) T'(j, k), j <= i: rows i0..i0+3 per thread
    const int nib = (n + 3) >> 2;
    for (int w = tid; w < n * nib; w += NT) {
      const int j = w % n, i0 = (w / n) * 4;
      if (j > i0 + 3) continue;
      const double* c0 = C + (size_t)i0 * n;
      const double* c1 = C + (size_t)min(i0 + 1, n - 1) * n;
      const double* c2 = C + (size_t)min(i0 + 2, n - 1) * n;
      const double* c3 = C + (size_t)min(i0 + 3, n - 1) * n;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll 4
      for (int k = 0; k < n; ++k) {
        const double t = Tt[j + (size_t)k * n];
        s0 += c0[k] * t;
        s1 += c1[k] * t;
        s2 += c2[k] * t;
        s3 += c3[k] * t;
      }
      if (j <= i0) A[tri(i0) + j] = s0;
      if (i0 + 1 < n && j <= i0 + 1) A[tri(i0 + 1) + j] = s1;
      if (i0 + 2 < n && j <= i0 + 2) A[tri(i0 + 2) + j] = s2;
      if (i0 + 3 < n && j <= i0 + 3) A[tri(i0 + 3) + j] = s3;
    }
  } else {
    for (int e = tid; e < n * n; e += NT) {
      const int r = e % n, c = e / n;
      C[e] = (r == c) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  const int nblk = tri(npair);  // 2x2 blocks (k1 >= k2)
  // the blocks of this thread, k1 | k2 << 8, kept in registers
  constexpr int KB = NT >= 256 ? (kMaxBatchN / 2 * (kMaxBatchN / 2 + 1) / 2 + NT - 1) / NT : 1;
  int blk[KB];
  if constexpr (NT >= 256) {
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      const int b = tid + j * NT;
      const int k1 = b < nblk ? tri_row(b) : 0;
      blk[j] = b < nblk ? k1 | ((b - tri(k1)) << 8) : -1;
    }
  }
  // A <- R A R^T on the block (pair k1 rows, pair k2 columns), R = [c -s; s c]
  auto turn_block = [&](int k1, int k2) {
    const double2 a1 = csn[k1], a2 = csn[k2];
    if (a1.y == 0.0 && a2.y == 0.0) return;
    const double c1 = a1.x, s1 = a1.y, c2 = a2.x, s2 = a2.y;
    const int4 i1 = pq4[k1];
    if (k1 == k2) {
      const int ip = i1.z + i1.x, iq = i1.w + i1.y, ipq = i1.w + i1.x;
      const double x = A[ip], y = A[ipq], v = A[iq];
      // rows, then columns
      const double x1 = c1 * x - s1 * y, y1 = s1 * x + c1 * y;
      const double u1 = c1 * y - s1 * v, v1 = s1 * y + c1 * v;
      A[ip] = c1 * x1 - s1 * u1;
      A[iq] = s1 * y1 + c1 * v1;
      A[ipq] = 0.0;
    } else {
      const int4 i2 = pq4[k2];
      const int ixx = i1.x >= i2.x ? i1.z + i2.x : i2.z + i1.x;
      const int iyx = i1.y >= i2.x ? i1.w + i2.x : i2.z + i1.y;
      const int ixu = i1.x >= i2.y ? i1.z + i2.y : i2.w + i1.x;
      const int iyu = i1.y >= i2.y ? i1.w + i2.y : i2.w + i1.y;
      const double x = A[ixx], y = A[iyx], u = A[ixu], v = A[iyu];
      const double x1 = c1 * x - s1 * y, y1 = s1 * x + c1 * y;
      const double u1 = c1 * u - s1 * v, v1 = s1 * u + c1 * v;
      A[ixx] = c2 * x1 - s2 * u1;
      A[ixu] = s2 * x1 + c2 * u1;
      A[iyx] = c2 * y1 - s2 * v1;
      A[iyu] = s2 * y1 + c2 * v1;
    }
  };
  const int stepk = NT / n, stepr = NT % n;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < tri(n); e += NT) {
      const int i = tri_row(e);
      const double v = fabs(A[e]);
      if (e == tri(i) + i) dg = fmax(dg, v);
      else off = fmax(off, v);
    }
    if (tid == 0) rotated = 0;
    wg_max2<NT>(off, dg, red);
    const double lim = tol * fmax(dg, 1.0);
    if (off <= lim) break;
    const double skip = 1.0e-2 * lim;
    for (int r = 0; r < np - 1; ++r) {
      for (int k = tid; k < npair; k += NT) {
        int p = rr_pos(k, r, np), q = rr_pos(np - 1 - k, r, np);
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double app = A[tri(p) + p], aqq = A[tri(q) + q], apq = A[tri(q) + p];
          // |a_pq| > 1e-18 sqrt|a_pp a_qq| compared as squares
          if (fabs(apq) > skip && apq * apq > 1.0e-36 * fabs(app * aqq)) {
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = rsqrt(t * t + 1.0);
            s = t * c;
            atomicAdd(&nrot, 1);
          }
        }
        csn[k] = make_double2(c, s);
        pq4[k] = make_int4(p, q, tri(p), tri(q));
        if (!CLDS) {
          lg[2 * (r * npair + k)] = c;
          lg[2 * (r * npair + k) + 1] = s;
          if (s != 0.0) rotated = 1;
        }
      }
      __syncthreads();
      if constexpr (NT >= 256) {
#pragma unroll
        for (int j = 0; j < KB; ++j) {
          if (blk[j] < 0) break;
          turn_block(blk[j] & 0xff, blk[j] >> 8);
        }
      } else {
        for (int b = tid; b < nblk; b += NT) {
          const int k1 = tri_row(b);
          turn_block(k1, b - tri(k1));
        }
      }
      // C <- C R^T on columns (p, q)
      if (CLDS) {
        int k = tid / n, row = tid % n;
        for (int e = tid; e < npair * n; e += NT) {
          const double2 a = csn[k];
          if (a.y != 0.0) {
            const int4 iq = pq4[k];
            double* cp = C + iq.x * n + row;
            double* cq = C + iq.y * n + row;
            const double u = *cp, v = *cq;
            *cp = a.x * u - a.y * v;
            *cq = a.y * u + a.x * v;
          }
          k += stepk;
          row += stepr;
          if (row >= n) { row -= n; ++k; }
        }
      }
      __syncthreads();
    }
    // every thread reads the flag before thread 0 may clear it for the next sweep
    const bool pending = !CLDS && rotated != 0;
    __syncthreads();
    if (pending) {
      for (int e = tid; e < na; e += NT) park[e] = A[e];
      __syncthreads();
      beig_flush_c<NT>(C, lg, n, np, dyn, room, csn2, off2);
      for (int e = tid; e < na; e += NT) A[e] = park[e];
      __syncthreads();
    }
  }
  // ascending order, ties by index
  for (int i = tid; i < n; i += NT) {
    const double ai = A[tri(i) + i];
    int rk = 0;
    for (int j = 0; j < n; ++j) {
      const double aj = A[tri(j) + j];
      rk += (aj < ai || (aj == ai && j < i)) ? 1 : 0;
    }
    rank[i] = rk;
  }
  __syncthreads();
  double* sorted = CLDS ? M.C : Tt;
  for (int e = tid; e < n * n; e += NT) {
    const int r = e % n, c = e / n;
    sorted[r + rank[c] * n] = C[e];
  }
  for (int i = tid; i < n; i += NT) M.eig[rank[i]] = A[tri(i) + i];
  __syncthreads();
  if (!CLDS)
    for (int e = tid; e < n * n; e += NT) M.C[e] = Tt[e];
  __syncthreads();
  // P = 2 C_occ C_occ^T, packed, with C_occ staged in LDS where it fits
  const int nocc = M.nocc;
  const double* Cg = M.C;
  if (n * nocc <= room) {
    for (int e = tid; e < n * nocc; e += NT) dyn[e] = Cg[e];
    Cg = dyn;
    __syncthreads();
  }
  for (int q = tid; q < tri(n); q += NT) {
    const int r = tri_row(q), c = q - tri(r);
    double s = 0.0;
#pragma unroll 4
    for (int k = 0; k < nocc; ++k) s += Cg[r + k * n] * Cg[c + k * n];
    M.p[q] = 2.0 * s;
  }
  if (tid == 0) {
    st[m].eig_rot += nrot;
    st[m].eig_flops += beig_flops(n, M.nocc, nrot, warm);
    st[m].have_c = 1;
  }
}

// ---- control, part 3: PL, frontier levels, newdg (iter.F90:1030-1041) -------
__global__ void k_bctl_post(int nmol, const BatchMol* __restrict__ mols, BatchState* __restrict__ st,
                            const int* __restrict__ do_eig, const double* __restrict__ scal) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol || !do_eig[m]) return;
  BatchState s = st[m];
  const BatchMol& M = mols[m];
  s.pl = scal[2 * m];
  const int ih = max(M.nocc, 1);
  if (ih < M.n) {
    s.eig_homo = M.eig[ih - 1];
    s.eig_lumo = M.eig[ih];
  }
  const double trans = 0.2;
  s.oknewd = (s.pl < s.sellim) || s.oknewd;
  s.newdg = ((s.pl < trans) && s.oknewd) || s.newdg;
  if (s.pl < trans * 0.3333) s.oknewd = 1;
  st[m] = s;
}

// Pa = P / 2 for the molecules that stepped
__global__ __launch_bounds__(256) void k_bhalf(const BatchMol* __restrict__ mols, const int* __restrict__ blk_mol,
                                               const int* __restrict__ blk_row, const int* __restrict__ do_eig) {
  const int m = blk_mol[blockIdx.x];
  if (!do_eig[m]) return;
  const BatchMol& M = mols[m];
  const long long r = blk_row[blockIdx.x];
  for (long long c = threadIdx.x; c <= r; c += 256) {
    const long long q = r * (r + 1) / 2 + c;
    M.pa[q] = 0.5 * M.p[q];
  }
}

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace
}  // namespace mhip

using namespace mhip;

struct mhip_batch {
  int nmol = 0;
  hipStream_t s = nullptr;
  std::vector<int> n_h, nocc_h, numat_h;
  // device storage
  DevBuf w, h, f, p, pa, pold, p1, C, A, T, eig, jrow, jcol;
  DevBuf nf0, norb, woff, wcum, w1off, ccol, heavy, light, nhb, nlb;
  DevBuf args, mols, cnv, st, act, doeig, mode, niter, shift, ee, partial, scal, orboff;
  std::vector<double> p0, pa0, p1h;  // starting densities, restored by every run
  DevBuf atom_mol, atom_idx, row_mol, row_idx;
  int n_atom_blk = 0, n_row_blk = 0;
  // molecules by eigensolver class (workgroup size), in launch order
  std::vector<int> cls_count, cls_nmax;
  int eig_form[4] = {0, 0, 0, 0};  // 0 global, 1 A in LDS, 2 A and C in LDS
  size_t eig_lds[4] = {0, 0, 0, 0};
  bool force_global_eig = false;
  DevBuf cls_list;
  int tile[4] = {64, 128, 256, 256};   // workgroup size per class (n <= 32, 64, 128, 256)
  double tol = 1.0e-10;
  std::vector<BatchState> init_state;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  // eager timing of the eigensolver launches per size class
  hipEvent_t eig_ev[4][2] = {};
  bool eig_timing = false;
  int graph_iters = 0;
};

namespace {

int size_class(int n) { return n <= 32 ? 0 : (n <= 64 ? 1 : (n <= 128 ? 2 : 3)); }

// LDS bytes a workgroup can take beyond the kernel's static arrays
size_t lds_room(const void* fn) {
  int dev = 0, lds_max = 0;
  MHIP_CHECK(hipGetDevice(&dev));
  MHIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  hipFuncAttributes fa{};
  MHIP_CHECK(hipFuncGetAttributes(&fa, fn));
  return lds_max > (int)fa.sharedSizeBytes ? (size_t)lds_max - fa.sharedSizeBytes : 0;
}

// Eigensolver form of a size class (its largest n, workgroup size NT): A and
// C in LDS when both fit (2), else A in LDS (1), else the global-memory
// kernel (0); the LDS forms get their dynamic LDS limit raised here, before
// any capture.  Returns the form and sets the dynamic LDS bytes.
template <int NT>
int plan_beig(int nmax, size_t* lds) {
  const size_t np = (size_t)((nmax + 1) & ~1);
  const size_t a = sizeof(double) * (np * (np + 1) / 2), c = sizeof(double) * (size_t)nmax * nmax;
  const void* f2 = (const void*)k_beig_lds<NT, true>;
  const void* f1 = (const void*)k_beig_lds<NT, false>;
  if (a + c <= lds_room(f2)) {
    MHIP_CHECK(hipFuncSetAttribute(f2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(a + c)));
    *lds = a + c;
    return 2;
  }
  if (a <= lds_room(f1)) {
    MHIP_CHECK(hipFuncSetAttribute(f1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)a));
    *lds = a;
    return 1;
  }
  *lds = 0;
  return 0;
}

template <int NT>
void launch_beig(hipStream_t s, int form, size_t lds, int nblk, const BatchMol* mols, const int* doeig,
                 BatchState* st, const int* list, double tol) {
  const int room = (int)(lds / sizeof(double));
  if (form == 2)
    hipLaunchKernelGGL((k_beig_lds<NT, true>), dim3(nblk), dim3(NT), lds, s, mols, doeig, st, list, tol, room);
  else if (form == 1)
    hipLaunchKernelGGL((k_beig_lds<NT, false>), dim3(nblk), dim3(NT), lds, s, mols, doeig, st, list, tol, room);
  else
    hipLaunchKernelGGL(k_beig<NT>, dim3(nblk), dim3(NT), 0, s, mols, doeig, st, list, tol);
  MHIP_LAUNCH_CHECK();
}

void plan_eig_classes(mhip_batch* b) {
  for (int c = 0; c < 4; ++c) {
    size_t lds = 0;
    int form = 0;
    if (b->cls_nmax[c] > 0) {
      switch (b->tile[c]) {
        case 64: form = plan_beig<64>(b->cls_nmax[c], &lds); break;
        case 128: form = plan_beig<128>(b->cls_nmax[c], &lds); break;
        case 512: form = plan_beig<512>(b->cls_nmax[c], &lds); break;
        case 1024: form = plan_beig<1024>(b->cls_nmax[c], &lds); break;
        default: form = plan_beig<256>(b->cls_nmax[c], &lds); break;
      }
    }
    b->eig_form[c] = b->force_global_eig ? 0 : form;
    b->eig_lds[c] = b->force_global_eig ? 0 : lds;
  }
}

void launch_eig_class(mhip_batch* b, int cls, int offset, int count) {
  if (count <= 0) return;
  const int* list = b->cls_list.as<int>() + offset;
  auto* mols = b->mols.as<BatchMol>();
  auto* st = b->st.as<BatchState>();
  const int* de = b->doeig.as<int>();
  const int f = b->eig_form[cls];
  const size_t l = b->eig_lds[cls];
  switch (b->tile[cls]) {
    case 64: launch_beig<64>(b->s, f, l, count, mols, de, st, list, b->tol); break;
    case 128: launch_beig<128>(b->s, f, l, count, mols, de, st, list, b->tol); break;
    case 512: launch_beig<512>(b->s, f, l, count, mols, de, st, list, b->tol); break;
    case 1024: launch_beig<1024>(b->s, f, l, count, mols, de, st, list, b->tol); break;
    default: launch_beig<256>(b->s, f, l, count, mols, de, st, list, b->tol); break;
  }
}

void one_iteration(mhip_batch* b) {
  hipStream_t s = b->s;
  const int nm = b->nmol;
  const int cb = cdiv(nm, 64);
  hipLaunchKernelGGL(k_bctl_pre, dim3(cb), dim3(64), 0, s, nm, b->st.as<BatchState>(), b->act.as<int>(),
                     b->mode.as<int>(), b->niter.as<int>(), b->shift.as<double>());
  MHIP_LAUNCH_CHECK();
  fock_build_batch(s, b->args.as<FockArgs>(), b->n_atom_blk, b->atom_mol.as<int>(), b->atom_idx.as<int>(),
                   b->act.as<int>(), b->mode.as<int>(), b->shift.as<double>());
  helect_batch(s, b->args.as<FockArgs>(), nm, b->n_row_blk, b->row_mol.as<int>(), b->row_idx.as<int>(),
               b->act.as<int>(), b->orboff.as<long long>(), b->partial.as<double>(), 2.0, b->ee.as<double>());
  hipLaunchKernelGGL(k_bctl_mid, dim3(cb), dim3(64), 0, s, nm, b->st.as<BatchState>(), b->ee.as<double>(),
                     b->doeig.as<int>());
  MHIP_LAUNCH_CHECK();
  int off = 0;
  for (int c = 0; c < 4; ++c) {
    if (b->eig_timing) MHIP_CHECK(hipEventRecord(b->eig_ev[c][0], s));
    launch_eig_class(b, c, off, b->cls_count[c]);
    if (b->eig_timing) MHIP_CHECK(hipEventRecord(b->eig_ev[c][1], s));
    off += b->cls_count[c];
  }
  cnvg_batch(s, b->cnv.as<BatchCnvg>(), nm, b->n_row_blk, b->row_mol.as<int>(), b->row_idx.as<int>(),
             b->doeig.as<int>(), b->niter.as<int>(), b->scal.as<double>());
  hipLaunchKernelGGL(k_bctl_post, dim3(cb), dim3(64), 0, s, nm, b->mols.as<BatchMol>(), b->st.as<BatchState>(),
                     b->doeig.as<int>(), b->scal.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bhalf, dim3(b->n_row_blk), dim3(256), 0, s, b->mols.as<BatchMol>(), b->row_mol.as<int>(),
                     b->row_idx.as<int>(), b->doeig.as<int>());
  MHIP_LAUNCH_CHECK();
}

template <class F>
int bguard(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    return 1;
  }
}

}  // namespace

extern "C" {

int mhip_batch_create(int nmol, const mhip_batch_molecule* mols, mhip_batch** out) {
  return bguard([&] {
    MHIP_REQUIRE(nmol > 0 && mols && out, "mhip_batch_create: bad arguments");
    auto* b = new mhip_batch();
    b->nmol = nmol;
    MHIP_CHECK(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
    // ---- host tables, then one upload per array
    std::vector<long long> o_w(nmol), o_mp(nmol), o_nn(nmol), o_n(nmol), o_at(nmol), o_cc(nmol), o_a(nmol);
    long long tw = 0, tmp = 0, tnn = 0, tn = 0, tat = 0, tcc = 0, ta = 0;
    std::vector<int> nf0_all, norb_all, heavy_all, light_all, nhb_all, nlb_all;
    std::vector<long long> woff_all, wcum_all, w1off_all, ccol_all;
    for (int m = 0; m < nmol; ++m) {
      const mhip_batch_molecule& M = mols[m];
      const int n = M.norbs, na = M.numat;
      MHIP_REQUIRE(n > 0 && n <= kMaxBatchN && na > 0, "mhip_batch_create: molecule size out of range (n <= 256)");
      MHIP_REQUIRE(M.nclose >= 0 && M.nclose <= n, "mhip_batch_create: bad nclose");
      b->n_h.push_back(n);
      b->nocc_h.push_back(M.nclose);
      b->numat_h.push_back(na);
      o_w[m] = tw;
      o_mp[m] = tmp;
      o_nn[m] = tnn;
      o_n[m] = tn;
      o_at[m] = tat;
      o_cc[m] = tcc;
      // atom-pair tables (the same construction as the single-system upload)
      long long kr = 0, wc = 0, cc = 0;
      int nh = 0, nl = 0;
      for (int i = 0; i < na; ++i) {
        const int f0 = M.nfirst[i] - 1, no = M.nlast[i] - M.nfirst[i] + 1;
        MHIP_REQUIRE(no == 1 || no == 4, "mhip_batch_create: only s and sp atoms");
        MHIP_REQUIRE(i == 0 ? f0 == 0 : f0 == nf0_all.back() + norb_all.back(),
                     "mhip_batch_create: orbitals must be contiguous by atom");
        nf0_all.push_back(f0);
        norb_all.push_back(no);
        const int lim = no * (no + 1) / 2;
        woff_all.push_back(kr);
        wcum_all.push_back(wc);
        kr += (long long)lim * wc;
        w1off_all.push_back(kr);
        kr += (long long)lim * lim;
        wc += lim;
        nhb_all.push_back(nh);
        nlb_all.push_back(nl);
        if (no == 4) { heavy_all.push_back(i); ++nh; } else { light_all.push_back(i); ++nl; }
        ccol_all.push_back(cc);
        cc += (long long)(na - 1 - i) * lim;
      }
      MHIP_REQUIRE(kr == M.n2elec, "mhip_batch_create: n2elec does not match the hcore layout");
      MHIP_REQUIRE(nf0_all.back() + norb_all.back() == n, "mhip_batch_create: nlast(numat) != norbs");
      tw += kr;
      tmp += (long long)n * (n + 1) / 2;
      tnn += (long long)n * n;
      {
        // the eigensolver's A: n x n (global form), or the LDS form's
        // rotation log of one sweep and parked A
        const long long np = (n + 1) & ~1;
        o_a[m] = ta;
        ta += std::max((long long)n * n, np * (np - 1) + np * (np + 1) / 2);
      }
      tn += n;
      tat += na;
      tcc += std::max(cc, 1LL);
    }
    // heavy/light lists hold indices within each molecule: offsets per molecule
    std::vector<long long> o_heavy(nmol), o_light(nmol);
    {
      long long hh = 0, ll = 0;
      for (int m = 0; m < nmol; ++m) {
        o_heavy[m] = hh;
        o_light[m] = ll;
        for (int i = 0; i < mols[m].numat; ++i) {
          if (mols[m].nlast[i] - mols[m].nfirst[i] == 3) ++hh; else ++ll;
        }
      }
    }
    hipStream_t s = b->s;
    b->w.ensure(sizeof(double) * std::max(tw, 1LL));
    b->h.ensure(sizeof(double) * tmp);
    for (DevBuf* x : {&b->f, &b->p, &b->pa, &b->pold}) x->ensure(sizeof(double) * tmp);
    b->p1.ensure(sizeof(double) * tn);
    b->eig.ensure(sizeof(double) * tn);
    for (DevBuf* x : {&b->C, &b->T}) x->ensure(sizeof(double) * tnn);
    b->A.ensure(sizeof(double) * ta);
    b->jrow.ensure(sizeof(double) * 10 * tat);
    b->jcol.ensure(sizeof(double) * tcc);
    for (int m = 0; m < nmol; ++m) {
      const mhip_batch_molecule& M = mols[m];
      const long long mp = (long long)M.norbs * (M.norbs + 1) / 2;
      if (M.n2elec)
        MHIP_CHECK(hipMemcpyAsync(b->w.as<double>() + o_w[m], M.w, sizeof(double) * M.n2elec,
                                  hipMemcpyHostToDevice, s));
      MHIP_CHECK(hipMemcpyAsync(b->h.as<double>() + o_mp[m], M.h, sizeof(double) * mp, hipMemcpyHostToDevice, s));
    }
    auto upv = [&](DevBuf& d, const void* src, size_t bytes) {
      d.ensure(bytes ? bytes : 8);
      if (bytes) MHIP_CHECK(hipMemcpyAsync(d.ptr, src, bytes, hipMemcpyHostToDevice, s));
    };
    upv(b->nf0, nf0_all.data(), sizeof(int) * nf0_all.size());
    upv(b->norb, norb_all.data(), sizeof(int) * norb_all.size());
    upv(b->woff, woff_all.data(), sizeof(long long) * woff_all.size());
    upv(b->wcum, wcum_all.data(), sizeof(long long) * wcum_all.size());
    upv(b->w1off, w1off_all.data(), sizeof(long long) * w1off_all.size());
    upv(b->ccol, ccol_all.data(), sizeof(long long) * ccol_all.size());
    upv(b->heavy, heavy_all.data(), sizeof(int) * heavy_all.size());
    upv(b->light, light_all.data(), sizeof(int) * light_all.size());
    upv(b->nhb, nhb_all.data(), sizeof(int) * nhb_all.size());
    upv(b->nlb, nlb_all.data(), sizeof(int) * nlb_all.size());
    // per-molecule argument blocks
    std::vector<FockArgs> args(nmol);
    std::vector<BatchMol> bm(nmol);
    std::vector<BatchCnvg> cv(nmol);
    std::vector<long long> orboff(nmol);
    std::vector<int> atom_mol, atom_idx, row_mol, row_idx;
    b->init_state.resize(nmol);
    // initial densities (iter.F90:269-300 with pdiag)
    std::vector<double> p0(tmp, 0.0), pa0(tmp, 0.0), p1h(tn, 0.0);
    for (int m = 0; m < nmol; ++m) {
      const mhip_batch_molecule& M = mols[m];
      const int n = M.norbs;
      FockArgs& a = args[m];
      a.sys.numat = M.numat;
      a.sys.norbs = n;
      a.sys.n2elec = M.n2elec;
      a.sys.nf0 = b->nf0.as<int>() + o_at[m];
      a.sys.norb = b->norb.as<int>() + o_at[m];
      a.sys.woff = b->woff.as<long long>() + o_at[m];
      a.sys.wcum = b->wcum.as<long long>() + o_at[m];
      a.sys.w1off = b->w1off.as<long long>() + o_at[m];
      a.sys.ccol = b->ccol.as<long long>() + o_at[m];
      a.sys.heavy = b->heavy.as<int>() + o_heavy[m];
      a.sys.light = b->light.as<int>() + o_light[m];
      a.sys.nheavy_before = b->nhb.as<int>() + o_at[m];
      a.sys.nlight_before = b->nlb.as<int>() + o_at[m];
      a.sys.w = b->w.as<double>() + o_w[m];
      a.h = b->h.as<double>() + o_mp[m];
      a.p = b->p.as<double>() + o_mp[m];
      a.pa = b->pa.as<double>() + o_mp[m];
      a.fin = b->f.as<double>() + o_mp[m];
      a.f = b->f.as<double>() + o_mp[m];
      a.jrow = b->jrow.as<double>() + 10 * o_at[m];
      a.jcol = b->jcol.as<double>() + o_cc[m];
      BatchMol& q = bm[m];
      q.n = n;
      q.nocc = M.nclose;
      q.numat = M.numat;
      q.f = a.f;
      q.p = b->p.as<double>() + o_mp[m];
      q.pa = b->pa.as<double>() + o_mp[m];
      q.pold = b->pold.as<double>() + o_mp[m];
      q.p1 = b->p1.as<double>() + o_n[m];
      q.C = b->C.as<double>() + o_nn[m];
      q.A = b->A.as<double>() + o_a[m];
      q.T = b->T.as<double>() + o_nn[m];
      q.eig = b->eig.as<double>() + o_n[m];
      cv[m] = BatchCnvg{n, q.p, q.pold, q.p1};
      orboff[m] = o_n[m];
      for (int i = 0; i < M.numat; ++i) {
        atom_mol.push_back(m);
        atom_idx.push_back(i);
      }
      for (int r = 0; r < n; ++r) {
        row_mol.push_back(m);
        row_idx.push_back(r);
      }
      const double w1 = M.nclose / (M.nclose + 1.0e-6 + M.nclose);
      for (int i = 0; i < n; ++i) {
        const long long d = o_mp[m] + (long long)i * (i + 1) / 2 + i;
        p0[d] = M.pdiag[i];
        pa0[d] = M.pdiag[i] * (w1 < 1e-6 ? 0.5 : w1);
        p1h[o_n[m] + i] = pa0[d];  // pold2 = diagonal of pold (= Pa)
      }
      BatchState& z = b->init_state[m];
      z = BatchState{};
      z.itrmax = M.itrmax > 0 ? M.itrmax : 2000;
      z.ten = -80.0;
      z.shift = 1.0;
      z.shfmax = 20.0;
      z.pl = 1.0;
      z.eold = 100.0;
      z.iredy = 1;
      z.enuclr = M.enuclr;
      z.atheat = M.atheat;
      const double scf = std::max(M.scfcrt > 0 ? M.scfcrt : 1.0e-4, 1e-12);
      z.selcon = std::min(std::min(scf * 100.0, 0.1), scf);
      z.pltest = 0.05 * std::sqrt(std::fabs(z.selcon));
    }
    b->n_atom_blk = (int)atom_mol.size();
    b->n_row_blk = (int)row_mol.size();
    upv(b->args, args.data(), sizeof(FockArgs) * nmol);
    upv(b->mols, bm.data(), sizeof(BatchMol) * nmol);
    upv(b->cnv, cv.data(), sizeof(BatchCnvg) * nmol);
    upv(b->orboff, orboff.data(), sizeof(long long) * nmol);
    upv(b->atom_mol, atom_mol.data(), sizeof(int) * atom_mol.size());
    upv(b->atom_idx, atom_idx.data(), sizeof(int) * atom_idx.size());
    upv(b->row_mol, row_mol.data(), sizeof(int) * row_mol.size());
    upv(b->row_idx, row_idx.data(), sizeof(int) * row_idx.size());
    b->st.ensure(sizeof(BatchState) * nmol);
    for (DevBuf* x : {&b->act, &b->doeig, &b->mode, &b->niter}) x->ensure(sizeof(int) * nmol);
    for (DevBuf* x : {&b->shift, &b->ee}) x->ensure(sizeof(double) * nmol);
    b->scal.ensure(sizeof(double) * 2 * nmol);
    b->partial.ensure(sizeof(double) * tn);
    // the starting P, Pa, Pold, p1, uploaded by every mhip_batch_run
    b->p0 = std::move(p0);
    b->pa0 = std::move(pa0);
    b->p1h = std::move(p1h);
    // molecules grouped by eigensolver class
    std::vector<int> order;
    b->cls_count.assign(4, 0);
    b->cls_nmax.assign(4, 0);
    for (int c = 0; c < 4; ++c)
      for (int m = 0; m < nmol; ++m)
        if (size_class(b->n_h[m]) == c) {
          order.push_back(m);
          b->cls_count[c]++;
          b->cls_nmax[c] = std::max(b->cls_nmax[c], b->n_h[m]);
        }
    upv(b->cls_list, order.data(), sizeof(int) * nmol);
    plan_eig_classes(b);
    MHIP_CHECK(hipStreamSynchronize(s));
    *out = b;
  });
}

int mhip_batch_set_tiles(mhip_batch* b, const int* tiles) {
  return bguard([&] {
    MHIP_REQUIRE(b && tiles, "mhip_batch_set_tiles: bad arguments");
    for (int c = 0; c < 4; ++c) {
      MHIP_REQUIRE(tiles[c] == 64 || tiles[c] == 128 || tiles[c] == 256 || tiles[c] == 512 || tiles[c] == 1024,
                   "mhip_batch_set_tiles: workgroup sizes are 64, 128, 256, 512 or 1024");
      b->tile[c] = tiles[c];
    }
    plan_eig_classes(b);
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
    }
  });
}

int mhip_batch_set_eig_lds(mhip_batch* b, int lds) {
  return bguard([&] {
    MHIP_REQUIRE(b, "mhip_batch_set_eig_lds: bad arguments");
    b->force_global_eig = lds == 0;
    plan_eig_classes(b);
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
    }
  });
}

// the per-molecule state and the starting densities, before a run
static void reset_batch(mhip_batch* b) {
  hipStream_t s = b->s;
  MHIP_CHECK(hipMemcpyAsync(b->st.ptr, b->init_state.data(), sizeof(BatchState) * b->nmol,
                            hipMemcpyHostToDevice, s));
  MHIP_CHECK(hipMemcpyAsync(b->p.ptr, b->p0.data(), sizeof(double) * b->p0.size(), hipMemcpyHostToDevice, s));
  MHIP_CHECK(hipMemcpyAsync(b->pa.ptr, b->pa0.data(), sizeof(double) * b->pa0.size(), hipMemcpyHostToDevice, s));
  MHIP_CHECK(hipMemcpyAsync(b->pold.ptr, b->pa0.data(), sizeof(double) * b->pa0.size(), hipMemcpyHostToDevice, s));
  MHIP_CHECK(hipMemcpyAsync(b->p1.ptr, b->p1h.data(), sizeof(double) * b->p1h.size(), hipMemcpyHostToDevice, s));
  MHIP_CHECK(hipStreamSynchronize(s));
}

int mhip_batch_eig_profile(mhip_batch* b, int iters, double* ms, double* flops, long long* rotations) {
  return bguard([&] {
    MHIP_REQUIRE(b && iters > 0 && ms && flops, "mhip_batch_eig_profile: bad arguments");
    hipStream_t s = b->s;
    reset_batch(b);
    for (int c = 0; c < 4; ++c) {
      ms[c] = 0.0;
      flops[c] = 0.0;
      if (rotations) rotations[c] = 0;
      for (int e = 0; e < 2; ++e) MHIP_CHECK(hipEventCreate(&b->eig_ev[c][e]));
    }
    b->eig_timing = true;
    struct Off {
      mhip_batch* b;
      ~Off() {
        b->eig_timing = false;
        for (auto& c : b->eig_ev)
          for (auto& e : c)
            if (e) {
              (void)hipEventDestroy(e);
              e = nullptr;
            }
      }
    } off{b};
    for (int it = 0; it < iters; ++it) {
      one_iteration(b);
      MHIP_CHECK(hipStreamSynchronize(s));
      for (int c = 0; c < 4; ++c) {
        float t = 0.0f;
        MHIP_CHECK(hipEventElapsedTime(&t, b->eig_ev[c][0], b->eig_ev[c][1]));
        ms[c] += t;
      }
    }
    std::vector<BatchState> hs(b->nmol);
    MHIP_CHECK(hipMemcpy(hs.data(), b->st.ptr, sizeof(BatchState) * b->nmol, hipMemcpyDeviceToHost));
    for (int m = 0; m < b->nmol; ++m) {
      const int c = size_class(b->n_h[m]);
      flops[c] += hs[m].eig_flops;
      if (rotations) rotations[c] += hs[m].eig_rot;
    }
  });
}

// Run the batched SCF from the starting densities: capture `iters`
// iterations once into a graph, replay it until every molecule is done.
int mhip_batch_run(mhip_batch* b, int iters_per_graph, int max_launches, double* seconds, int* n_converged) {
  return bguard([&] {
    MHIP_REQUIRE(b && iters_per_graph > 0 && max_launches > 0, "mhip_batch_run: bad arguments");
    hipStream_t s = b->s;
    reset_batch(b);
    if (!b->exec || b->graph_iters != iters_per_graph) {
      if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
      }
      MHIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (int it = 0; it < iters_per_graph; ++it) one_iteration(b);
      MHIP_CHECK(hipStreamEndCapture(s, &b->graph));
      MHIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_iters = iters_per_graph;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<BatchState> hs(b->nmol);
    int done = 0;
    for (int l = 0; l < max_launches; ++l) {
      MHIP_CHECK(hipGraphLaunch(b->exec, s));
      MHIP_CHECK(hipMemcpyAsync(hs.data(), b->st.ptr, sizeof(BatchState) * b->nmol, hipMemcpyDeviceToHost, s));
      MHIP_CHECK(hipStreamSynchronize(s));
      done = 0;
      for (const auto& z : hs) done += z.done ? 1 : 0;
      if (done == b->nmol) break;
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (n_converged) {
      int c = 0;
      for (const auto& z : hs) c += z.converged ? 1 : 0;
      *n_converged = c;
    }
  });
}

int mhip_batch_results(mhip_batch* b, double* escf, int* niter, int* converged) {
  return bguard([&] {
    std::vector<BatchState> hs(b->nmol);
    MHIP_CHECK(hipMemcpy(hs.data(), b->st.ptr, sizeof(BatchState) * b->nmol, hipMemcpyDeviceToHost));
    for (int m = 0; m < b->nmol; ++m) {
      if (escf) escf[m] = hs[m].escf;
      if (niter) niter[m] = hs[m].niter;
      if (converged) converged[m] = hs[m].converged;
    }
  });
}

int mhip_batch_density(mhip_batch* b, int m, double* p) {
  return bguard([&] {
    MHIP_REQUIRE(b && m >= 0 && m < b->nmol && p, "mhip_batch_density: bad arguments");
    long long off = 0;
    for (int k = 0; k < m; ++k) off += (long long)b->n_h[k] * (b->n_h[k] + 1) / 2;
    const long long mp = (long long)b->n_h[m] * (b->n_h[m] + 1) / 2;
    MHIP_CHECK(hipMemcpy(p, b->p.as<double>() + off, sizeof(double) * mp, hipMemcpyDeviceToHost));
  });
}

void mhip_batch_destroy(mhip_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->s);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (DevBuf* x : {&b->w, &b->h, &b->f, &b->p, &b->pa, &b->pold, &b->p1, &b->C, &b->A, &b->T, &b->eig, &b->jrow,
                    &b->jcol, &b->nf0, &b->norb, &b->woff, &b->wcum, &b->w1off, &b->ccol, &b->heavy, &b->light,
                    &b->nhb, &b->nlb, &b->args, &b->mols, &b->cnv, &b->st, &b->act, &b->doeig, &b->mode, &b->niter,
                    &b->shift, &b->ee, &b->partial, &b->scal, &b->orboff, &b->atom_mol, &b->atom_idx,
                    &b->row_mol, &b->row_idx, &b->cls_list})
    x->release();
  (void)hipStreamDestroy(b->s);
  delete b;
}

}  // extern "C"
