// Two-electron Fock assembly over atom-pair blocks, the electronic energy and
// the density extrapolation (cnvg), all on MOPAC packed storage.
//
// Fock build (reference: src/SCF/fock2.F90:149-369 with jab.F90, kab.F90,
// fock1.F90) as two passes with no atomics:
//   row pass    one workgroup per atom i streams its contiguous W row
//               (pairs (i, j<i), hcore order), reduces the Coulomb field on
//               atom i over all partners, writes the exchange block F(i,j)
//               (unique to the pair) and parks the Coulomb contribution to
//               atom j in a column-ordered scratch;
//   column pass one workgroup per atom j sums its parked contributions (a
//               contiguous range), adds the row-pass sum and the one-centre
//               terms and writes the diagonal atom block of F.
// Partners are visited heavy atoms first, then hydrogens, so a wavefront
// never mixes the 10x10 and 10x1 block shapes.
#include "common.h"
#include "k_fock.h"

namespace mhip {
namespace {

__constant__ int c_pa[10] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3};  // larger orbital of pair
__constant__ int c_pb[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};  // smaller orbital of pair

__device__ __forceinline__ long long pk(long long r, long long c) {
  return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r;
}

// Starting value of F(q) before G(P) is added.
__device__ __forceinline__ double finit(const FockArgs& a, long long q, bool diag) {
  switch (a.init_mode) {
    case kInitShift: {
      double v = a.h[q] + a.shift * a.pa[q] + 1.0e-16 * (double)(q + 1);
      if (diag) v -= a.shift;
      return v;
    }
    case kInitRandom:
      return a.h[q] + (((q + 1) & 1) ? -0.001 : 0.001);
    case kInitCopy:
      return a.fin[q];
    default:
      return a.h[q];
  }
}

// Row pass for atom i (one NT-thread workgroup); shared by the single-system
// (256 threads) and the batched (64 threads: a small molecule's atom has few
// partners) launches.
template <int NT>
__device__ void fock_rows_body(const FockArgs& a, const int i) {
  const int ni = a.sys.norb[i];
  const long long fi = a.sys.nf0[i];
  const double* __restrict__ P = a.p;
  const double* __restrict__ Pa = a.pa;
  // density of atom i with off-diagonal elements doubled
  double pi2[10];
  if (ni == 4) {
#pragma unroll
    for (int ab = 0; ab < 10; ++ab) {
      const int x = c_pa[ab], y = c_pb[ab];
      pi2[ab] = P[pk(fi + x, fi + y)] * (x == y ? 1.0 : 2.0);
    }
  } else {
    pi2[0] = P[pk(fi, fi)];
#pragma unroll
    for (int ab = 1; ab < 10; ++ab) pi2[ab] = 0.0;
  }
  double ji[10];
#pragma unroll
  for (int ab = 0; ab < 10; ++ab) ji[ab] = 0.0;

  const long long wrow = a.sys.woff[i];
  const int limi = (ni == 4) ? 10 : 1;
  const int nheavy = a.sys.nheavy_before[i];
  const int nlight = a.sys.nlight_before[i];

  // ---- heavy partners -------------------------------------------------------
  for (int t = threadIdx.x; t < nheavy; t += NT) {
    const int j = a.sys.heavy[t];
    const long long fj = a.sys.nf0[j];
    double pj2[10];
#pragma unroll
    for (int cd = 0; cd < 10; ++cd) {
      const int x = c_pa[cd], y = c_pb[cd];
      pj2[cd] = P[pk(fj + x, fj + y)] * (x == y ? 1.0 : 2.0);
    }
    const double* __restrict__ wb = a.sys.w + wrow + (long long)limi * a.sys.wcum[j];
    double jj[10];
#pragma unroll
    for (int cd = 0; cd < 10; ++cd) jj[cd] = 0.0;
    const long long col = a.sys.ccol[j] + (long long)(i - j - 1) * 10;
    if (ni == 4) {
      double pa_ij[4][4], K[4][4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          pa_ij[x][y] = Pa[pk(fi + x, fj + y)];
          K[x][y] = 0.0;
        }
#pragma unroll
      for (int ab = 0; ab < 10; ++ab) {
        const int x = c_pa[ab], y = c_pb[ab];
        double wr[10];
#pragma unroll
        for (int cd = 0; cd < 10; ++cd) wr[cd] = wb[ab * 10 + cd];
        double s = 0.0;
#pragma unroll
        for (int cd = 0; cd < 10; ++cd) {
          s = fma(wr[cd], pj2[cd], s);
          jj[cd] = fma(pi2[ab], wr[cd], jj[cd]);
          const int u = c_pa[cd], v = c_pb[cd];
          const double wv = wr[cd];
          K[x][u] = fma(pa_ij[y][v], wv, K[x][u]);
          if (x != y) K[y][u] = fma(pa_ij[x][v], wv, K[y][u]);
          if (u != v) K[x][v] = fma(pa_ij[y][u], wv, K[x][v]);
          if (x != y && u != v) K[y][v] = fma(pa_ij[x][u], wv, K[y][v]);
        }
        ji[ab] += s;
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          const long long q = pk(fi + x, fj + y);
          a.f[q] = finit(a, q, false) - K[x][y];
        }
    } else {
      // hydrogen i with heavy j: W block is the 10 pairs of atom j
      const double pii = pi2[0];
      double pa_ij[4], K[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int y = 0; y < 4; ++y) pa_ij[y] = Pa[pk(fi, fj + y)];
      double s = 0.0;
#pragma unroll
      for (int cd = 0; cd < 10; ++cd) {
        const double wv = wb[cd];
        s = fma(wv, pj2[cd], s);
        jj[cd] = pii * wv;
        const int u = c_pa[cd], v = c_pb[cd];
        K[u] = fma(pa_ij[v], wv, K[u]);
        if (u != v) K[v] = fma(pa_ij[u], wv, K[v]);
      }
      ji[0] += s;
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const long long q = pk(fi, fj + y);
        a.f[q] = finit(a, q, false) - K[y];
      }
    }
#pragma unroll
    for (int cd = 0; cd < 10; ++cd) a.jcol[col + cd] = jj[cd];
  }

  // ---- hydrogen partners ----------------------------------------------------
  for (int t = threadIdx.x; t < nlight; t += NT) {
    const int j = a.sys.light[t];
    const long long fj = a.sys.nf0[j];
    const double pjj = P[pk(fj, fj)];
    const double* __restrict__ wb = a.sys.w + wrow + (long long)limi * a.sys.wcum[j];
    const long long col = a.sys.ccol[j] + (long long)(i - j - 1);
    if (ni == 4) {
      double pa_ij[4], K[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int x = 0; x < 4; ++x) pa_ij[x] = Pa[pk(fi + x, fj)];
      double s = 0.0;
#pragma unroll
      for (int ab = 0; ab < 10; ++ab) {
        const double wv = wb[ab];
        ji[ab] = fma(wv, pjj, ji[ab]);
        s = fma(pi2[ab], wv, s);
        const int x = c_pa[ab], y = c_pb[ab];
        K[x] = fma(pa_ij[y], wv, K[x]);
        if (x != y) K[y] = fma(pa_ij[x], wv, K[y]);
      }
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const long long q = pk(fi + x, fj);
        a.f[q] = finit(a, q, false) - K[x];
      }
      a.jcol[col] = s;
    } else {
      const double wv = wb[0];
      ji[0] = fma(wv, pjj, ji[0]);
      const long long q = pk(fi, fj);
      a.f[q] = finit(a, q, false) - Pa[q] * wv;
      a.jcol[col] = wv * pi2[0];
    }
  }

  // ---- reduce the Coulomb field on atom i over the workgroup -----------------
  if constexpr (NT == 64) {
    // one wave: lane exchange, fixed order
#pragma unroll
    for (int ab = 0; ab < 10; ++ab)
      for (int o = 32; o > 0; o >>= 1) ji[ab] += __shfl_xor(ji[ab], o);
    if (threadIdx.x == 0)
      for (int ab = 0; ab < limi; ++ab) a.jrow[(long long)i * 10 + ab] = ji[ab];
  } else {
    __shared__ double red[10][NT];
#pragma unroll
    for (int ab = 0; ab < 10; ++ab) red[ab][threadIdx.x] = ji[ab];
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w)
#pragma unroll
        for (int ab = 0; ab < 10; ++ab) red[ab][threadIdx.x] += red[ab][threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x < limi) a.jrow[(long long)i * 10 + threadIdx.x] = red[threadIdx.x][0];
  }
}

// atoms in descending order: atom i has i partners, so the longest rows
// start first and the short ones fill in behind them
__global__ __launch_bounds__(256) void k_fock_rows(FockArgs a) {
  fock_rows_body<256>(a, a.sys.numat - 1 - (int)blockIdx.x);
}

// Column pass for atom j (one 256-thread workgroup).
__device__ void fock_cols_body(const FockArgs& a, const int j) {
  const int nj = a.sys.norb[j];
  const int lim = (nj == 4) ? 10 : 1;
  const long long fj = a.sys.nf0[j];
  const long long m = (long long)(a.sys.numat - 1 - j) * lim;
  const double* __restrict__ src = a.jcol + a.sys.ccol[j];
  const int stride = (lim == 10) ? 250 : 256;
  double s = 0.0;
  if (threadIdx.x < stride)
    for (long long idx = threadIdx.x; idx < m; idx += stride) s += src[idx];
  __shared__ double red[256];
  red[threadIdx.x] = (threadIdx.x < stride) ? s : 0.0;
  __syncthreads();
  if (threadIdx.x < lim) {
    // sum the partial sums that belong to this pair index (tid % lim == cd)
    double t = 0.0;
    for (int k = threadIdx.x; k < stride; k += lim) t += red[k];
    const int cd = threadIdx.x;
    const int x = c_pa[cd], y = c_pb[cd];
    // one-centre terms (fock1.F90:38-88)
    const double* __restrict__ w1 = a.sys.w + a.sys.w1off[j];
    double g = 0.0;
    const int iw = x + 1, jw = y + 1;
    const int ijw = iw * (iw - 1) / 2 + jw;
    for (int k = 0; k < nj; ++k) {
      const int kw = k + 1;
      for (int l = 0; l < nj; ++l) {
        const int lw = l + 1;
        const long long q = pk(fj + k, fj + l);
        int im = kw > lw ? kw : lw, jm = kw > lw ? lw : kw;
        const int klw = im * (im - 1) / 2 + jm;
        im = kw > jw ? kw : jw; jm = kw > jw ? jw : kw;
        const int ikw = im * (im - 1) / 2 + jm;
        im = lw > iw ? lw : iw; jm = lw > iw ? iw : lw;
        const int jlw = im * (im - 1) / 2 + jm;
        g += a.p[q] * w1[(klw - 1) * lim + (ijw - 1)] - a.pa[q] * w1[(jlw - 1) * lim + (ikw - 1)];
      }
    }
    const long long q = pk(fj + x, fj + y);
    a.f[q] = finit(a, q, x == y) + a.jrow[(long long)j * 10 + cd] + t + g;
  }
}

__global__ __launch_bounds__(256) void k_fock_cols(FockArgs a) { fock_cols_body(a, blockIdx.x); }

// ---- electronic energy -------------------------------------------------------
__global__ __launch_bounds__(256) void k_helect_rows(int n, const double* __restrict__ p,
                                                     const double* __restrict__ h,
                                                     const double* __restrict__ f,
                                                     double* __restrict__ partial) {
  const long long r = blockIdx.x;
  const long long base = r * (r + 1) / 2;
  double s = 0.0;
  for (long long c = threadIdx.x; c <= r; c += 256) {
    const long long q = base + c;
    const double v = p[q] * (h[q] + f[q]);
    s += (c == r) ? 0.5 * v : v;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[r] = red[0];
}

__global__ __launch_bounds__(1024) void k_sum(int n, const double* __restrict__ x, double scale,
                                              double* __restrict__ out) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = scale * red[0];
}

// ---- cnvg (src/SCF/cnvg.F90:16-145), RHF --------------------------------------
// One workgroup does every diagonal step (statistics, damping, clamp and the
// renormalization loop); the off-diagonal extrapolation only needs `fac`.
template <int NTH>
__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = NTH / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}
template <int NTH>
__device__ double block_max(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = NTH / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

template <int NTH>
__device__ void cnvg_diag_body(int n, double* __restrict__ pnew, double* __restrict__ pold,
                               double* __restrict__ p1, int niter, double* __restrict__ scal) {
  __shared__ double red[NTH];
  const double rhfuhf = 2.0;
  const double damp = (niter > 3) ? 0.05 : 1.0e10;
  const bool extrap = (niter % 3) != 0;
  double sum1 = 0.0, pl = 0.0, faca = 0.0, facb = 0.0;
  for (int i = threadIdx.x; i < n; i += NTH) {
    const long long k = (long long)i * (i + 1) / 2 + i;
    const double a = pnew[k];
    sum1 += a;
    const double sa = fabs(a - pold[k]);
    pl = fmax(pl, sa);
    if (!extrap) {
      faca += sa * sa;
      const double t = a - 2.0 * pold[k] + p1[i];
      facb += t * t;
    }
    p1[i] = pold[k];
    pold[k] = a;
  }
  sum1 = block_sum<NTH>(sum1, red);
  pl = block_max<NTH>(pl, red);
  faca = block_sum<NTH>(faca, red);
  facb = block_sum<NTH>(facb, red);
  double fac = 0.0;
  if (facb > 1.0e-10 && faca < 100.0 * facb) fac = sqrt(faca / facb);
  double sum2 = 0.0;
  for (int i = threadIdx.x; i < n; i += NTH) {
    const long long k = (long long)i * (i + 1) / 2 + i;
    double v = pold[k];
    if (fabs(v - p1[i]) > damp) v = p1[i] + ((v - p1[i]) >= 0.0 ? damp : -damp);
    else v = v + fac * (v - p1[i]);
    v = fmin(rhfuhf, fmax(v, 0.0));
    pold[k] = v;
    pnew[k] = v;
    sum2 += v;
  }
  sum2 = block_sum<NTH>(sum2, red);
  const double sum0 = sum1;
  for (int it = 0; it < 1000; ++it) {
    const double sc = (sum2 > 1.0e-3) ? sum1 / sum2 : 0.0;
    sum1 = sum0;
    if (sum2 < 1.0e-3 || fabs(sc - 1.0) < 1.0e-5) break;
    double s2 = 0.0, nclamp = 0.0;
    for (int i = threadIdx.x; i < n; i += NTH) {
      const long long k = (long long)i * (i + 1) / 2 + i;
      double v = pold[k] * sc + 1.0e-20;
      v = fmax(v, 0.0);
      if (v > rhfuhf) {
        v = rhfuhf;
        nclamp += 1.0;
      } else {
        s2 += v;
      }
      pold[k] = v;
      pnew[k] = v;
    }
    sum2 = block_sum<NTH>(s2, red);
    sum1 -= rhfuhf * block_sum<NTH>(nclamp, red);
  }
  if (threadIdx.x == 0) {
    scal[0] = pl;
    scal[1] = fac;
  }
}

__global__ __launch_bounds__(1024) void k_cnvg_diag(int n, double* __restrict__ pnew,
                                                    double* __restrict__ pold,
                                                    double* __restrict__ p1, int niter,
                                                    double* __restrict__ scal) {
  cnvg_diag_body<1024>(n, pnew, pold, p1, niter, scal);
}

__global__ void k_cnvg_offdiag(int n, double* __restrict__ pnew, double* __restrict__ pold,
                               const double* __restrict__ scal) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (r >= n || c >= r) return;
  const double fac = scal[1];
  const long long q = (long long)r * (r + 1) / 2 + c;
  const double a = pnew[q];
  const double v = a + fac * (a - pold[q]);
  pold[q] = v;
  pnew[q] = v;
}

// ---- batched forms: one launch for many molecules --------------------------
// blk_mol / blk_idx map each workgroup to (molecule, atom or row); molecules
// whose flag in `active` is 0 (converged) do nothing.
__global__ __launch_bounds__(64) void k_bfock_rows(const FockArgs* __restrict__ args,
                                                    const int* __restrict__ blk_mol,
                                                    const int* __restrict__ blk_idx,
                                                    const int* __restrict__ active,
                                                    const int* __restrict__ mode,
                                                    const double* __restrict__ shift) {
  const int m = blk_mol[blockIdx.x];
  if (!active[m]) return;
  FockArgs a = args[m];
  a.init_mode = mode[m];
  a.shift = shift[m];
  fock_rows_body<64>(a, blk_idx[blockIdx.x]);
}

__global__ __launch_bounds__(256) void k_bfock_cols(const FockArgs* __restrict__ args,
                                                    const int* __restrict__ blk_mol,
                                                    const int* __restrict__ blk_idx,
                                                    const int* __restrict__ active,
                                                    const int* __restrict__ mode,
                                                    const double* __restrict__ shift) {
  const int m = blk_mol[blockIdx.x];
  if (!active[m]) return;
  FockArgs a = args[m];
  a.init_mode = mode[m];
  a.shift = shift[m];
  fock_cols_body(a, blk_idx[blockIdx.x]);
}

// electronic energy: rows (blk_mol, blk_idx = row), then one sum per molecule
__global__ __launch_bounds__(256) void k_bhelect_rows(const FockArgs* __restrict__ args,
                                                      const int* __restrict__ blk_mol,
                                                      const int* __restrict__ blk_idx,
                                                      const int* __restrict__ active,
                                                      const long long* __restrict__ orb_off,
                                                      double* __restrict__ partial) {
  const int m = blk_mol[blockIdx.x];
  if (!active[m]) return;
  const FockArgs& a = args[m];
  const long long r = blk_idx[blockIdx.x];
  const long long base = r * (r + 1) / 2;
  double s = 0.0;
  for (long long c = threadIdx.x; c <= r; c += 256) {
    const long long q = base + c;
    const double v = a.pa[q] * (a.h[q] + a.f[q]);
    s += (c == r) ? 0.5 * v : v;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[orb_off[m] + r] = red[0];
}

__global__ __launch_bounds__(256) void k_bhelect_sum(const FockArgs* __restrict__ args,
                                                     const int* __restrict__ active,
                                                     const long long* __restrict__ orb_off,
                                                     const double* __restrict__ partial, double scale,
                                                     double* __restrict__ ee) {
  const int m = blockIdx.x;
  if (!active[m]) return;
  const int n = args[m].sys.norbs;
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[orb_off[m] + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) ee[m] = scale * red[0];
}

// cnvg: the diagonal pass per molecule, then the off-diagonal rows
__global__ __launch_bounds__(256) void k_bcnvg_diag(const BatchCnvg* __restrict__ b,
                                                    const int* __restrict__ active,
                                                    const int* __restrict__ niter,
                                                    double* __restrict__ scal) {
  const int m = blockIdx.x;
  if (!active[m]) return;
  const BatchCnvg& c = b[m];
  cnvg_diag_body<256>(c.n, c.p, c.pold, c.p1, niter[m], scal + 2 * m);
}

__global__ __launch_bounds__(256) void k_bcnvg_offdiag(const BatchCnvg* __restrict__ b,
                                                       const int* __restrict__ blk_mol,
                                                       const int* __restrict__ blk_idx,
                                                       const int* __restrict__ active,
                                                       const double* __restrict__ scal) {
  const int m = blk_mol[blockIdx.x];
  if (!active[m]) return;
  const BatchCnvg& c = b[m];
  const int r = blk_idx[blockIdx.x];
  const double fac = scal[2 * m + 1];
  for (int col = threadIdx.x; col < r; col += 256) {
    const long long q = (long long)r * (r + 1) / 2 + col;
    const double a = c.p[q];
    const double v = a + fac * (a - c.pold[q]);
    c.pold[q] = v;
    c.p[q] = v;
  }
}

__global__ void k_axpby_packed(long long len, double a, const double* __restrict__ x, double b,
                               double* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < len;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = (b == 0.0) ? a * x[i] : a * x[i] + b * y[i];
}

}  // namespace

void fock_build(hipStream_t s, const FockArgs& a) {
  // algorithmic traffic: W once, plus ptot, p, f in and f out (packed)
  const double mp = (double)a.sys.norbs * (a.sys.norbs + 1) / 2;
  timed(s, K_FOCK, 0.0, 8.0 * ((double)a.sys.n2elec + 4.0 * mp), [&] {
    hipLaunchKernelGGL(k_fock_rows, dim3(a.sys.numat), dim3(256), 0, s, a);
    MHIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_fock_cols, dim3(a.sys.numat), dim3(256), 0, s, a);
    MHIP_LAUNCH_CHECK();
  });
}

void helect(hipStream_t s, int n, const double* pa, const double* h, const double* f,
            double scale, double* partial, double* out) {
  hipLaunchKernelGGL(k_helect_rows, dim3(n), dim3(256), 0, s, n, pa, h, f, partial);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, n, partial, scale, out);
  MHIP_LAUNCH_CHECK();
}

void cnvg(hipStream_t s, int n, double* pnew, double* pold, double* p1, int niter, double* scal) {
  hipLaunchKernelGGL(k_cnvg_diag, dim3(1), dim3(1024), 0, s, n, pnew, pold, p1, niter, scal);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cnvg_offdiag, dim3((n + 255) / 256, n), dim3(256), 0, s, n, pnew, pold, scal);
  MHIP_LAUNCH_CHECK();
}

void fock_build_batch(hipStream_t s, const FockArgs* args, int nblk, const int* blk_mol, const int* blk_atom,
                      const int* active, const int* mode, const double* shift) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_bfock_rows, dim3(nblk), dim3(64), 0, s, args, blk_mol, blk_atom, active, mode, shift);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bfock_cols, dim3(nblk), dim3(256), 0, s, args, blk_mol, blk_atom, active, mode, shift);
  MHIP_LAUNCH_CHECK();
}

void helect_batch(hipStream_t s, const FockArgs* args, int nmol, int nrow_blk, const int* blk_mol,
                  const int* blk_row, const int* active, const long long* orb_off, double* partial,
                  double scale, double* ee) {
  if (nrow_blk <= 0) return;
  hipLaunchKernelGGL(k_bhelect_rows, dim3(nrow_blk), dim3(256), 0, s, args, blk_mol, blk_row, active, orb_off,
                     partial);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bhelect_sum, dim3(nmol), dim3(256), 0, s, args, active, orb_off, partial, scale, ee);
  MHIP_LAUNCH_CHECK();
}

void cnvg_batch(hipStream_t s, const BatchCnvg* b, int nmol, int nrow_blk, const int* blk_mol,
                const int* blk_row, const int* active, const int* niter, double* scal) {
  if (nmol <= 0) return;
  hipLaunchKernelGGL(k_bcnvg_diag, dim3(nmol), dim3(256), 0, s, b, active, niter, scal);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bcnvg_offdiag, dim3(nrow_blk), dim3(256), 0, s, b, blk_mol, blk_row, active, scal);
  MHIP_LAUNCH_CHECK();
}

void axpby(hipStream_t s, long long len, double a, const double* x, double b, double* y) {
  const int blocks = (int)std::min<long long>(4096, (len + 255) / 256);
  hipLaunchKernelGGL(k_axpby_packed, dim3(blocks), dim3(256), 0, s, len, a, x, b, y);
  MHIP_LAUNCH_CHECK();
}

}  // namespace mhip
