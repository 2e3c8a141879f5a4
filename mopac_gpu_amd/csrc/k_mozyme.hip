// MOZYME localized-MO kernels: density_for_MOZYME, diagg1, diagg2
// (src/MOZYME/density_for_MOZYME.F90:16-342, diagg1.F90:16-649,
// diagg2.F90:16-359), on the reference's own LMO arrays.
//
// Layout (MOZYME_C.F90:116-162): LMO i's atoms are icocc[nncf[i] + k]
// (1-based atom numbers), its coefficients cocc[ncocc[i] + ...] atom after
// atom; the same for the virtual set (nce, nnce, icvir, ncvir, cvir).  The
// sparse P and F hold atom-pair blocks at nijbo(a, b) (0-based, -1 if absent;
// a > b: iorbs(a) x iorbs(b), b fastest; a == b: packed lower triangle).
//
// density: one workgroup per occupied LMO adds 2 c c^T into every stored
//   block of its atom pairs (FP64 atomics; the order of the additions is the
//   only difference from the reference, <= 1e-15 relative).
// diagg1: one workgroup per virtual LMO builds its F-weighted vector ws on
//   its own atoms in LDS, then scans all occupied LMOs with the reference's
//   fast test and cutoffs.  Two passes (count, then write at the prefix-sum
//   offset) keep the reference's list order (virtual LMO, then occupied LMO)
//   and its storage cap; a third kernel evaluates the occupied LMO energies.
// diagg2: the annihilation list is processed in the reference's order by one
//   workgroup (its rotations and the growth of the LMOs depend on each other
//   through the LMOs they share); inside a rotation the atoms of both LMOs
//   are handled in parallel, and growth keeps the reference's atom order by
//   a prefix scan.
#include <cmath>
#include <vector>

#include "common.h"
#include <cstring>
#include <map>

namespace mhip {
namespace {

struct LmoDev {
  int numat, norbs, nocc, nvir;
  const int* nfirst;  // 1-based
  const int* iorbs;
  const int* nijbo;   // column-major numat x numat
  int* ncf;
  const int* nncf;
  const int* ncocc;
  int* icocc;
  int icocc_dim;
  double* cocc;
  int cocc_dim;
  int* nce;
  const int* nnce;
  const int* ncvir;
  int* icvir;
  int icvir_dim;
  double* cvir;
  int cvir_dim;
};

__device__ __forceinline__ int nij_of(const LmoDev& L, int a1, int b1) {  // 1-based atoms
  return L.nijbo[(size_t)(a1 - 1) + (size_t)(b1 - 1) * L.numat];
}

// ---------------------------------------------------------------- density
__global__ void k_mz_density_init(long long n, int mode, const double* __restrict__ partp,
                                   double* __restrict__ p) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = (mode == 0) ? 0.0 : (mode == -1 ? -0.5 * partp[i] : 0.5 * partp[i]);
}

__global__ void k_scale(long long n, double s, double* __restrict__ p) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] *= s;
}

// The density (density_for_MOZYME.F90:16-342) as a gather, in the reference's
// order: per atom the occupied LMOs that hold it, in LMO order; then one
// thread per stored atom pair adds c_i(a) c_i(b)^T over the LMOs the two
// atoms share, in LMO order, onto the starting value -- the order in which
// density_for_MOZYME.F90 adds them, LMO after LMO -- and applies the spin
// factor.  No atomics on P, and P is the same from run to run.
__global__ void k_mz_fill_u64(long long n, unsigned long long* __restrict__ x) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = 0ull;
}

// bit i of bits[a * nw + i / 64]: occupied LMO i holds atom a
__global__ void k_mz_atom_bits(LmoDev L, int nw, unsigned long long* __restrict__ bits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.nocc) return;
  const int b0 = L.nncf[i], n = L.ncf[i];
  const unsigned long long bit = 1ull << (i & 63);
  for (int k = 0; k < n; ++k) atomicOr(bits + (size_t)(L.icocc[b0 + k] - 1) * nw + (i >> 6), bit);
}

// per atom: LMOs before each 64-LMO word (cum) and in all (cnt)
__global__ void k_mz_atom_cum(int numat, int nw, const unsigned long long* __restrict__ bits,
                              int* __restrict__ cum, int* __restrict__ cnt) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= numat) return;
  int c = 0;
  for (int w = 0; w < nw; ++w) {
    cum[(size_t)a * nw + w] = c;
    c += __popcll(bits[(size_t)a * nw + w]);
  }
  cnt[a] = c;
}

// exclusive scan of cnt[0..n) into off[0..n], one 1024-thread workgroup
__global__ __launch_bounds__(1024) void k_mz_scan_counts(int n, const int* __restrict__ cnt, int* __restrict__ off) {
  __shared__ int part[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += 1024) {
    const int i = c0 + threadIdx.x;
    const int v = i < n ? cnt[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int t = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n) off[i] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) off[n] = carry;
}

// each (LMO, atom) entry straight to its place in the atom's list, in LMO
// order: the rank of LMO i among the atom's LMOs
__global__ void k_mz_atom_fill(LmoDev L, int nw, const unsigned long long* __restrict__ bits,
                               const int* __restrict__ cum, const int* __restrict__ off,
                               int* __restrict__ ent_lmo, int* __restrict__ ent_coff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.nocc) return;
  const int b0 = L.nncf[i], n = L.ncf[i];
  const unsigned long long below = (1ull << (i & 63)) - 1ull;
  int coff = L.ncocc[i];
  for (int k = 0; k < n; ++k) {
    const int a = L.icocc[b0 + k] - 1;
    const size_t wa = (size_t)a * nw + (i >> 6);
    const int slot = off[a] + cum[wa] + __popcll(bits[wa] & below);
    ent_lmo[slot] = i;
    ent_coff[slot] = coff;
    coff += L.iorbs[a];
  }
}

__global__ void k_mz_density_gather(LmoDev L, int npairs, const int2* __restrict__ pairs,
                                    const int* __restrict__ off, const int* __restrict__ ent_lmo,
                                    const int* __restrict__ ent_coff, int mode,
                                    const double* __restrict__ partp, double spinfa, int scale,
                                    double* __restrict__ p) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs) return;
  const int a = pairs[t].x, b = pairs[t].y;  // 0-based, a >= b
  const long long l0 = L.nijbo[(size_t)a + (size_t)b * L.numat];
  const int na = L.iorbs[a], nb = L.iorbs[b];
  const bool diag = a == b;
  // element (x, y) of the block: packed lower for the diagonal block,
  // row-major na x nb otherwise
  auto idx = [&](int x, int y) { return diag ? x * (x + 1) / 2 + y : x * nb + y; };
  double acc[16];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      double v = 0.0;
      if (x < na && y < nb && (!diag || y <= x) && mode != 0) {
        const double q = partp[l0 + idx(x, y)];
        v = mode == -1 ? -0.5 * q : 0.5 * q;
      }
      acc[x * 4 + y] = v;
    }
  auto add = [&](const double* ca, const double* cb) {
    double u[4], w[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      u[x] = x < na ? ca[x] : 0.0;
      w[x] = x < nb ? cb[x] : 0.0;
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y)
        if (x < na && y < nb && (!diag || y <= x)) acc[x * 4 + y] += u[x] * w[y];
  };
  int ia = off[a];
  const int ea = off[a + 1];
  if (diag) {
    for (; ia < ea; ++ia) {
      const double* c = L.cocc + ent_coff[ia];
      add(c, c);
    }
  } else {
    int ib = off[b];
    const int eb = off[b + 1];
    while (ia < ea && ib < eb) {
      const int la = ent_lmo[ia], lb = ent_lmo[ib];
      if (la < lb) {
        ++ia;
      } else if (lb < la) {
        ++ib;
      } else {
        add(L.cocc + ent_coff[ia], L.cocc + ent_coff[ib]);
        ++ia;
        ++ib;
      }
    }
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
      if (x < na && y < nb && (!diag || y <= x))
        p[l0 + idx(x, y)] = scale ? acc[x * 4 + y] * spinfa : acc[x * 4 + y];
}

// ---------------------------------------------------------------- diagg1
constexpr int MZ_MAXAT = 256;  // atoms per LMO handled in LDS

struct D1Params {
  double cutoff, flim, oldlim;
  int full;        // idiagg <= 5 or even: fast test + new list; else reuse the old list
  int fmo_dim;
  int fp32;        // the elements' products and sums in FP32 (the mixed form)
};

// Per virtual LMO: ws on its atoms (position-indexed), aov, avir, eigv.
// Shared by the count and the write passes.
struct VirtLds {
  int atoms[MZ_MAXAT];
  int woff[MZ_MAXAT];      // offset of the atom's orbitals in ws
  double avir[MZ_MAXAT];
  double aov[MZ_MAXAT];
  double ws[4 * MZ_MAXAT];
  double red[256];
  int nat;
};

// position of atom a (1-based) in the virtual LMO, or -1
__device__ __forceinline__ int vpos(const VirtLds& V, int a) {
  for (int k = 0; k < V.nat; ++k)
    if (V.atoms[k] == a) return k;
  return -1;
}

template <class T>
__device__ void ws_accumulate(const LmoDev& L, VirtLds& V, const double* __restrict__ fao,
                              const double* __restrict__ p, double cutoff, int loopi);

__device__ void virt_setup(const LmoDev& L, int i, const double* __restrict__ fao,
                           const double* __restrict__ p, double cutoff, VirtLds& V, double* eigv,
                           bool fp32) {
  const int tid = threadIdx.x;
  const int loopi = L.ncvir[i];
  if (tid == 0) {
    V.nat = min(L.nce[i], MZ_MAXAT);
    int o = 0;
    for (int k = 0; k < V.nat; ++k) {
      const int a = L.icvir[L.nnce[i] + k];
      V.atoms[k] = a;
      V.woff[k] = o;
      o += L.iorbs[a - 1];
    }
  }
  __syncthreads();
  // avir: squared weight of each atom in this virtual LMO
  for (int k = tid; k < V.nat; k += blockDim.x) {
    const int no = L.iorbs[V.atoms[k] - 1];
    double s = 0.0;
    for (int x = 0; x < no; ++x) {
      const double c = L.cvir[loopi + V.woff[k] + x];
      s += c * c;
    }
    V.avir[k] = s;
  }
  __syncthreads();
  // ws(j1 orbitals) = sum over atoms k1 of the LMO of F(j1, k1) c_vir(k1), as
  // diagg1.F90:142-195 accumulates it (k1 in LMO order); in FP32 for the
  // mixed form
  if (fp32) {
    ws_accumulate<float>(L, V, fao, p, cutoff, loopi);
  } else {
    ws_accumulate<double>(L, V, fao, p, cutoff, loopi);
  }
  __syncthreads();
  for (int k = tid; k < V.nat; k += blockDim.x) {
    const int no = L.iorbs[V.atoms[k] - 1];
    double s = 0.0;
    for (int x = 0; x < no; ++x) s += V.ws[V.woff[k] + x] * V.ws[V.woff[k] + x];
    V.aov[k] = s;
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int k = 0; k < V.nat; ++k) {
      if (V.aov[k] * V.avir[k] > cutoff) {
        const int no = L.iorbs[V.atoms[k] - 1];
        for (int x = 0; x < no; ++x) s += V.ws[V.woff[k] + x] * L.cvir[loopi + V.woff[k] + x];
      }
    }
    if (eigv) eigv[i] = s;
  }
  __syncthreads();
}

template <class T>
__device__ void ws_accumulate(const LmoDev& L, VirtLds& V, const double* __restrict__ fao,
                              const double* __restrict__ p, double cutoff, int loopi) {
  for (int jj = threadIdx.x; jj < V.nat; jj += blockDim.x) {
    const int j1 = V.atoms[jj], nj = L.iorbs[j1 - 1];
    T w[4] = {T(0), T(0), T(0), T(0)};
    for (int kk = 0; kk < V.nat; ++kk) {
      const int k1 = V.atoms[kk], nk = L.iorbs[k1 - 1];
      const int kj = nij_of(L, k1, j1);
      if (kj < 0) continue;
      if (!(V.avir[kk] * p[kj] > cutoff)) continue;
      const double* cv = L.cvir + loopi + V.woff[kk];
      if (nk == 1 && nj == 1) {
        w[0] += T(fao[kj]) * T(cv[0]);
      } else if (k1 > j1) {
        int ii = kj;
        for (int i4 = 0; i4 < nk; ++i4)
          for (int jx = 0; jx < nj; ++jx) w[jx] += T(fao[ii++]) * T(cv[i4]);
      } else if (k1 < j1) {
        int ii = kj;
        for (int jx = 0; jx < nj; ++jx)
          for (int i4 = 0; i4 < nk; ++i4) w[jx] += T(fao[ii++]) * T(cv[i4]);
      } else {
        for (int jx = 1; jx <= nj; ++jx)
          for (int i4 = 1; i4 <= nj; ++i4) {
            const int ii = (i4 > jx) ? kj + (i4 * (i4 - 1)) / 2 + jx : kj + (jx * (jx - 1)) / 2 + i4;
            w[jx - 1] += T(fao[ii - 1]) * T(cv[i4 - 1]);
          }
      }
    }
    for (int x = 0; x < nj; ++x) V.ws[V.woff[jj] + x] = (double)w[x];
  }
}

// Occupied LMO j against the virtual LMO in V: the fast test (full pass only)
// and the sum.  Returns 0 if the pair is not evaluated, 1 if evaluated but
// not stored, 2 if stored; value in *sum.
__device__ int occ_virt_element(const LmoDev& L, const VirtLds& V, const double* __restrict__ fao,
                                const double* __restrict__ aocc, int j, const D1Params& P,
                                bool fast_test, double* sum_out, const short* amap = nullptr) {
  if (fast_test) {
    const int i1 = V.atoms[0], i2 = (V.nat > 1) ? V.atoms[1] : i1;
    const int j1 = L.icocc[L.nncf[j]], j2 = (L.ncf[j] > 1) ? L.icocc[L.nncf[j] + 1] : j1;
    const int a = nij_of(L, i1, j1), b = nij_of(L, i1, j2), c = nij_of(L, i2, j1), d = nij_of(L, i2, j2);
    if (a < 0 && b < 0 && c < 0 && d < 0) return 0;
    double s = 0.0;
    if (a >= 0) s = fabs(fao[a]);
    if (b >= 0) s += fabs(fao[b]);
    if (c >= 0) s += fabs(fao[c]);
    if (d >= 0) s += fabs(fao[d]);
    if (!(s >= P.flim)) return 0;
  }
  const int loopj = L.ncocc[j];
  bool lij = false;
  double sum = 0.0;
  float sum32 = 0.0f;
  int kl = 0;
  for (int kk = L.nncf[j]; kk < L.nncf[j] + L.ncf[j]; ++kk) {
    const int k1 = L.icocc[kk], nk = L.iorbs[k1 - 1];
    const int vp = amap ? amap[k1 - 1] : vpos(V, k1);
    if (vp < 0 || aocc[kk] * V.aov[vp] < P.cutoff) {
      kl += nk;
    } else {
      lij = true;
      if (P.fp32) {
        for (int x = 0; x < nk; ++x) sum32 += (float)V.ws[V.woff[vp] + x] * (float)L.cocc[loopj + kl + x];
      } else {
        for (int x = 0; x < nk; ++x) sum += V.ws[V.woff[vp] + x] * L.cocc[loopj + kl + x];
      }
      kl += nk;
    }
  }
  if (P.fp32) sum = (double)sum32;
  *sum_out = sum;
  if (!fast_test) return 2;  // the reuse path stores every listed element
  return (lij && fabs(sum) > P.oldlim) ? 2 : 1;
}

// aocc[kk]: squared weight of the kk-th atom slot of the occupied LMOs
__global__ void k_mz_aocc(LmoDev L, double* __restrict__ aocc) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L.nocc) return;
  const int loopj = L.ncocc[j];
  int kl = 0;
  for (int kk = L.nncf[j]; kk < L.nncf[j] + L.ncf[j]; ++kk) {
    const int no = L.iorbs[L.icocc[kk] - 1];
    double s = 0.0;
    for (int x = 0; x < no; ++x) s += L.cocc[loopj + kl + x] * L.cocc[loopj + kl + x];
    aocc[kk] = s;
    kl += no;
  }
}

// Candidates of the full pass.  The fast test evaluates the pair (virtual
// i, occupied j) only if one of the atom pairs (i1 or i2, j1 or j2) -- the
// first two atoms of each LMO -- is stored.  So j can pass only if j1 or j2
// is a stored neighbour of i1 or i2: the neighbours come from nijbo (fixed
// per system), the occupied LMOs by first and second atom from per-atom
// lists built per call, and every other j is skipped exactly where the fast
// test would have returned 0.  The marks live in an LDS bitmap.
constexpr int MZ_CANDW = 256;  // bitmap words: up to 16384 occupied LMOs
struct CandDev {
  const int* nbr_off;  // numat + 1
  const int* nbr;      // stored neighbours (0-based), the atom itself included
  const int* fl_off;   // numat + 1
  const int* fl_ent;   // occupied LMOs whose first or second atom is this one
  int on;
};

__global__ void k_mz_zero_int(int n, int* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

__global__ void k_mz_first_count(LmoDev L, int* __restrict__ cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L.nocc) return;
  const int b0 = L.nncf[j];
  const int j1 = L.icocc[b0] - 1;
  atomicAdd(cnt + j1, 1);
  if (L.ncf[j] > 1) {
    const int j2 = L.icocc[b0 + 1] - 1;
    if (j2 != j1) atomicAdd(cnt + j2, 1);
  }
}

__global__ void k_mz_first_fill(LmoDev L, const int* __restrict__ off, int* __restrict__ fillc,
                                int* __restrict__ ent) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L.nocc) return;
  const int b0 = L.nncf[j];
  const int j1 = L.icocc[b0] - 1;
  ent[off[j1] + atomicAdd(fillc + j1, 1)] = j;
  if (L.ncf[j] > 1) {
    const int j2 = L.icocc[b0 + 1] - 1;
    if (j2 != j1) ent[off[j2] + atomicAdd(fillc + j2, 1)] = j;
  }
}

__device__ void mark_candidates(const LmoDev& L, const VirtLds& V, const CandDev& C,
                                unsigned long long* __restrict__ cand) {
  for (int w = threadIdx.x; w < MZ_CANDW; w += blockDim.x) cand[w] = 0ull;
  __syncthreads();
  const int i1 = V.atoms[0] - 1, i2 = (V.nat > 1 ? V.atoms[1] : V.atoms[0]) - 1;
  for (int h = 0; h < 2; ++h) {
    const int x = h == 0 ? i1 : i2;
    if (h == 1 && i2 == i1) break;
    for (int e = C.nbr_off[x] + threadIdx.x; e < C.nbr_off[x + 1]; e += blockDim.x) {
      const int b = C.nbr[e];
      for (int f = C.fl_off[b]; f < C.fl_off[b + 1]; ++f) {
        const int j = C.fl_ent[f];
        atomicOr(cand + (j >> 6), 1ull << (j & 63));
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ bool is_candidate(const CandDev& C, const unsigned long long* cand, int j) {
  return !C.on || ((cand[j >> 6] >> (j & 63)) & 1ull);
}

// atom (0-based) -> its position in the virtual LMO, or -1: replaces the
// list search of vpos for systems of up to MZ_AMAP atoms
constexpr int MZ_AMAP = 8192;
__device__ const short* atom_map(const LmoDev& L, const VirtLds& V, short* amap) {
  if (L.numat > MZ_AMAP) return nullptr;
  for (int a = threadIdx.x; a < L.numat; a += blockDim.x) amap[a] = -1;
  __syncthreads();
  for (int k = threadIdx.x; k < V.nat; k += blockDim.x) amap[V.atoms[k] - 1] = (short)k;
  __syncthreads();
  return amap;
}

// Pass 1 (full lists): per virtual LMO, the number of stored elements.
__global__ __launch_bounds__(256) void k_mz_diagg1_count(LmoDev L, const double* __restrict__ fao,
                                                         const double* __restrict__ p,
                                                         const double* __restrict__ aocc,
                                                         D1Params P, CandDev C, int* __restrict__ counts,
                                                         double* __restrict__ eigv) {
  __shared__ VirtLds V;
  __shared__ unsigned long long cand[MZ_CANDW];
  __shared__ short amap_s[MZ_AMAP];
  const int i = blockIdx.x;
  virt_setup(L, i, fao, p, P.cutoff, V, eigv, P.fp32 != 0);
  if (C.on) mark_candidates(L, V, C, cand);
  const short* amap = atom_map(L, V, amap_s);
  int cnt = 0;
  for (int j = threadIdx.x; j < L.nocc; j += blockDim.x) {
    double s;
    if (is_candidate(C, cand, j) && occ_virt_element(L, V, fao, aocc, j, P, true, &s, amap) == 2) ++cnt;
  }
  // block sum
  __shared__ int red[256];
  red[threadIdx.x] = cnt;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[i] = red[0];
}

// Pass 2: write the list at the prefix offsets in the reference's order,
// honouring the storage cap exactly (a candidate is evaluated only while the
// list is not full); per-LMO maxima and sums of |element| for tiny and sumt.
__global__ __launch_bounds__(256) void k_mz_diagg1_write(LmoDev L, const double* __restrict__ fao,
                                                         const double* __restrict__ p,
                                                         const double* __restrict__ aocc,
                                                         D1Params P, CandDev C, const int* __restrict__ offs,
                                                         int* __restrict__ nfmo,
                                                         const int* __restrict__ old_ifmo,
                                                         double* __restrict__ fmo, int* __restrict__ ifmo,
                                                         double* __restrict__ vmax,
                                                         double* __restrict__ vsum,
                                                         double* __restrict__ eigv) {
  __shared__ VirtLds V;
  __shared__ int scan[256];
  __shared__ int base_s;
  __shared__ unsigned long long cand[MZ_CANDW];
  __shared__ short amap_s[MZ_AMAP];
  const int i = blockIdx.x, tid = threadIdx.x;
  virt_setup(L, i, fao, p, P.cutoff, V, eigv, P.fp32 != 0);
  double mx = 0.0, sm = 0.0;
  if (P.full && C.on) mark_candidates(L, V, C, cand);
  const short* amap = atom_map(L, V, amap_s);
  if (P.full) {
    if (tid == 0) base_s = offs[i];
    __syncthreads();
    for (int j0 = 0; j0 < L.nocc; j0 += blockDim.x) {
      const int base = base_s;
      if (base >= P.fmo_dim) break;  // list full: nothing more is evaluated
      const int j = j0 + tid;
      double s = 0.0;
      int kind = 0;
      if (j < L.nocc && is_candidate(C, cand, j)) kind = occ_virt_element(L, V, fao, aocc, j, P, true, &s, amap);
      // inclusive scan of stored flags in j order
      scan[tid] = (kind == 2);
      __syncthreads();
      for (int off = 1; off < 256; off <<= 1) {
        const int t = (tid >= off) ? scan[tid - off] : 0;
        __syncthreads();
        scan[tid] += t;
        __syncthreads();
      }
      const int before = base + scan[tid] - (kind == 2);  // entries stored before this j
      if (kind != 0 && before < P.fmo_dim) {
        mx = fmax(mx, fabs(s));
        sm += fabs(s);
        if (kind == 2) {
          fmo[before] = s;
          ifmo[2 * before] = i + 1;
          ifmo[2 * before + 1] = j + 1;
        }
      }
      __syncthreads();
      if (tid == 255) base_s = base + scan[255];
      __syncthreads();
    }
    if (tid == 0) nfmo[i] = max(0, min(base_s, P.fmo_dim) - offs[i]);
  } else {
    // reuse: the old list entries of this virtual LMO, recomputed in place
    const int start = offs[i], cnt = nfmo[i];
    for (int jj = tid; jj < cnt; jj += blockDim.x) {
      const int pos = start + jj;
      if (pos >= P.fmo_dim) break;
      const int j = old_ifmo[2 * pos + 1] - 1;
      double s;
      occ_virt_element(L, V, fao, aocc, j, P, false, &s, amap);
      fmo[pos] = s;
      ifmo[2 * pos] = i + 1;
      ifmo[2 * pos + 1] = j + 1;
      mx = fmax(mx, fabs(s));
      sm += fabs(s);
    }
  }
  V.red[tid] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) V.red[tid] = fmax(V.red[tid], V.red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) vmax[i] = V.red[0];
  __syncthreads();
  V.red[tid] = sm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) V.red[tid] += V.red[tid + w];
    __syncthreads();
  }
  if (tid == 0) vsum[i] = V.red[0];
}

// Occupied LMO energies (diagg1.F90:484-640); one wave per occupied LMO.
// The reference's sum runs over (atom j, atom k, orbital jx) in that order;
// lanes take 64 atoms k at a time and form each product c_j(jx) s1(k, jx)
// with the reference's own arithmetic (s1 summed over k's orbitals in its
// order), and lane 0 adds the products in the reference's order, so the
// energy rounds exactly as the sequential loop does.
__device__ __forceinline__ void mz_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64) void k_mz_occ_energy(LmoDev L, const double* __restrict__ fao,
                                                      const double* __restrict__ p,
                                                      const double* __restrict__ aocc, double cutoff,
                                                      double* __restrict__ eigs) {
  const int i = blockIdx.x, lane = threadIdx.x;
  if (i >= L.nocc) return;
  __shared__ double term[64 * 4];
  __shared__ int tok[64];
  const int loopi = L.ncocc[i], b0 = L.nncf[i], nat_i = L.ncf[i];
  double sum = 0.0;  // lane 0's running sum
  int jl = loopi;
  for (int j = b0; j < b0 + nat_i; ++j) {
    const int j1 = L.icocc[j], nj = L.iorbs[j1 - 1];
    int carry = loopi;  // coefficient offset of the chunk's first atom
    for (int c0 = 0; c0 < nat_i; c0 += 64) {
      const bool in = c0 + lane < nat_i;
      const int k = b0 + c0 + lane;
      const int k1 = in ? L.icocc[k] : 1;
      const int nk = in ? L.iorbs[k1 - 1] : 0;
      // exclusive prefix of the orbital counts over the chunk
      int incl = nk;
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
      }
      const int kl = carry + incl - nk;
      carry += __shfl(incl, 63);
      int ok = 0;
      if (in) {
        const int k1j1 = nij_of(L, k1, j1);
        // avir(k1) of the reference here is the squared weight of atom k1 in LMO i
        if (k1j1 >= 0) {
          double wk = 0.0;
          for (int x = 0; x < nk; ++x) wk += L.cocc[kl + x] * L.cocc[kl + x];
          if (wk * p[k1j1] * aocc[k] >= cutoff) {
            ok = 1;
            for (int jx = 1; jx <= nj; ++jx) {
              double s1 = 0.0;
              if (k1 > j1) {
                for (int i4 = 1; i4 <= nk; ++i4) s1 += fao[k1j1 + (i4 - 1) * nj + jx - 1] * L.cocc[kl + i4 - 1];
              } else if (k1 < j1) {
                for (int i4 = 1; i4 <= nk; ++i4) s1 += fao[k1j1 + (jx - 1) * nk + i4 - 1] * L.cocc[kl + i4 - 1];
              } else {
                for (int j4 = 1; j4 <= jx; ++j4) s1 += fao[k1j1 + (jx * (jx - 1)) / 2 + j4 - 1] * L.cocc[kl + j4 - 1];
                for (int i4 = jx + 1; i4 <= nk; ++i4)
                  s1 += fao[k1j1 + (i4 * (i4 - 1)) / 2 + jx - 1] * L.cocc[kl + i4 - 1];
              }
              term[lane * 4 + jx - 1] = L.cocc[jl + jx - 1] * s1;
            }
          }
        }
      }
      tok[lane] = ok;
      mz_wave_sync();
      if (lane == 0) {
        const int cnt = min(64, nat_i - c0);
        for (int kk = 0; kk < cnt; ++kk)
          if (tok[kk])
            for (int jx = 0; jx < nj; ++jx) sum += term[kk * 4 + jx];
      }
      mz_wave_sync();
    }
    jl += nj;
  }
  if (lane == 0) eigs[i] = sum;
}

// ---------------------------------------------------------------- diagg2
struct D2Params {
  double tiny, biglim, cnst, shift, thresh;
  int retry, nij;
  int lcap;  // two LMOs' atoms at most this many: lists in LDS (<= D2W_LCAP)
  int pcap;  // coefficients of each LMO copied to LDS (<= D2W_PCAP; 0: global memory)
};

constexpr int D2_NT = 64;  // one wave: the per-rotation scans run at wave latency

// One rotation of the (virtual i, occupied j) list per workgroup.  A rotation
// reads and writes only LMOs i and j (coefficients, atom lists, counts), so
// rotations that share neither commute exactly: the host runs the list as
// levels (a rotation one level after the last earlier one on its i or its
// j), one launch per level, which reproduces the sequential order of
// diagg2.F90 for every LMO.  Each workgroup has its own slot of the atom
// flags (iused = -1, latoms = 0 between rotations) and of the LMO copies.
__global__ __launch_bounds__(D2_NT) void k_mz_diagg2(LmoDev L, const double* __restrict__ eigs,
                                                     const double* __restrict__ eigv,
                                                     const double* __restrict__ fmo,
                                                     const int* __restrict__ ifmo, D2Params P,
                                                     const int* __restrict__ order, int base,
                                                     int* __restrict__ iused_all,   // slots x numat, -1
                                                     char* __restrict__ latoms_all,  // slots x numat, 0
                                                     double* __restrict__ storej_all, int span_j,
                                                     double* __restrict__ storei_all, int span_i,
                                                     double* __restrict__ sumb_out,
                                                     int* __restrict__ nrej_out) {
  const int tid = threadIdx.x;
  const int slot = blockIdx.x;
  int* __restrict__ iused_g = iused_all + (size_t)slot * L.numat;
  char* __restrict__ latoms_g = latoms_all + (size_t)slot * L.numat;
  double* __restrict__ storej = storej_all + (size_t)slot * span_j;
  double* __restrict__ storei = storei_all + (size_t)slot * span_i;
  __shared__ int flag[D2_NT];
  __shared__ int s_mlf, s_mle, s_fail, s_ncf, s_nce, s_jext, s_iext;
  double sumb = 0.0;
  int nrej = 0;
  const int ij = order[base + slot];
  {
    const int i = ifmo[2 * ij] - 1, j = ifmo[2 * ij + 1] - 1;
    const double c = fmo[ij] * P.cnst;
    const double d = eigs[j] - eigv[i] - P.shift;
    const int ncfj = L.ncf[j], ncei = L.nce[i];
    const int jlr = L.ncocc[j];
    int jur, jncf, iur, incv;
    if (j != L.nocc - 1) { jur = L.ncocc[j + 1]; jncf = L.nncf[j + 1]; }
    else { jur = L.cocc_dim; jncf = L.icocc_dim; }
    jur = min(jlr + L.norbs, jur);  // exclusive end of LMO j's coefficients
    const int ilr = L.ncvir[i];
    if (i != L.nvir - 1) { iur = L.ncvir[i + 1]; incv = L.nnce[i + 1]; }
    else { iur = L.cvir_dim; incv = L.icvir_dim; }
    iur = min(ilr + L.norbs, iur);
    // undo log for a rejection: here the coefficients of the atoms the two
    // LMOs hold now (at most 4 per atom), later each grown position just
    // before it is written; a rejection puts back exactly what was written,
    // which leaves both stores as a full copy-and-restore would
    const int bj = min(4 * L.ncf[j], jur - jlr), bi = min(4 * L.nce[i], iur - ilr);
    for (int k = tid; k < bj; k += D2_NT) storej[k] = L.cocc[jlr + k];
    for (int k = tid; k < bi; k += D2_NT) storei[k] = L.cvir[ilr + k];
    const double e = copysign(sqrt(4.0 * c * c + d * d), d);
    double alpha = sqrt(0.5 * (1.0 + d / e));
    __syncthreads();
    for (;;) {
      const double beta = -copysign(sqrt(1.0 - alpha * alpha), c);
      if (tid == 0) sumb += fabs(beta);
      // iused: offset of each atom of occupied LMO j (its original atoms),
      // an exclusive scan of their orbital counts
      {
        const int n0 = L.nncf[j], nc = L.ncf[j];
        int base = 0;
        for (int c0 = 0; c0 < nc; c0 += D2_NT) {
          const int lf = c0 + tid;
          const int a = (lf < nc) ? L.icocc[n0 + lf] - 1 : -1;
          const int no = (a >= 0) ? L.iorbs[a] : 0;
          flag[tid] = no;
          __syncthreads();
          for (int off = 1; off < D2_NT; off <<= 1) {
            const int t = (tid >= off) ? flag[tid - off] : 0;
            __syncthreads();
            flag[tid] += t;
            __syncthreads();
          }
          if (a >= 0) iused_g[a] = base + flag[tid] - no;
          base += flag[D2_NT - 1];
          __syncthreads();
        }
        if (tid == 0) {
          s_mlf = base;
          s_fail = 0;
          s_ncf = L.ncf[j];
          s_nce = L.nce[i];
          s_jext = bj;
          s_iext = bi;
        }
      }
      __syncthreads();
      const int loopi = L.ncvir[i], loopj = L.ncocc[j];
      // ---- virtual LMO atoms, in order: rotate common atoms, grow LMO j
      {
        const int nv = L.nce[i];
        int mle_base = 0;  // running orbital offset over virtual atoms
        int mlf_run = s_mlf;
        int ncf_run = s_ncf;
        for (int c0 = 0; c0 < nv; c0 += D2_NT) {
          const int le = c0 + tid;
          int mie = -1, no = 0, grow = 0, mle = 0;
          // offsets of this chunk's atoms: exclusive scan of iorbs
          flag[tid] = (le < nv) ? L.iorbs[L.icvir[L.nnce[i] + le] - 1] : 0;
          __syncthreads();
          for (int off = 1; off < D2_NT; off <<= 1) {
            const int t = (tid >= off) ? flag[tid - off] : 0;
            __syncthreads();
            flag[tid] += t;
            __syncthreads();
          }
          if (le < nv) {
            mie = L.icvir[L.nnce[i] + le] - 1;
            no = L.iorbs[mie];
            mle = mle_base + flag[tid] - no;
          }
          const int chunk_orbs = flag[D2_NT - 1];
          __syncthreads();
          if (le < nv) {
            latoms_g[mie] = 1;
            const int u = iused_g[mie];
            if (u >= 0) {
              for (int x = 0; x < no; ++x) {
                const double a = L.cocc[loopj + u + x], b = L.cvir[loopi + mle + x];
                L.cocc[loopj + u + x] = alpha * a + beta * b;
                L.cvir[loopi + mle + x] = alpha * b - beta * a;
              }
            } else {
              double s = 0.0;
              for (int x = 0; x < no; ++x) {
                const double t = beta * L.cvir[loopi + mle + x];
                s += t * t;
              }
              grow = (s > P.thresh);
            }
          }
          // growth in order: scan of grow flags
          flag[tid] = grow;
          __syncthreads();
          for (int off = 1; off < D2_NT; off <<= 1) {
            const int t = (tid >= off) ? flag[tid - off] : 0;
            __syncthreads();
            flag[tid] += t;
            __syncthreads();
          }
          const int ngrow = flag[D2_NT - 1];
          if (ngrow > 0) {
            // orbital offsets of the grown atoms: second scan over their iorbs
            const int rank = flag[tid] - grow;  // grown atoms before this one
            __syncthreads();
            flag[tid] = grow ? no : 0;
            __syncthreads();
            for (int off = 1; off < D2_NT; off <<= 1) {
              const int t = (tid >= off) ? flag[tid - off] : 0;
              __syncthreads();
              flag[tid] += t;
              __syncthreads();
            }
            if (grow) {
              const int slot = L.nncf[j] + ncf_run + rank;
              const int obase = mlf_run + flag[tid] - no;
              // the reference stops at the first atom that would not fit
              if (slot >= jncf || jlr + obase + no > jur) {
                atomicOr(&s_fail, 1);
              }
            }
            __syncthreads();
            if (!s_fail && grow) {
              const int slot = L.nncf[j] + ncf_run + rank;
              const int obase = mlf_run + flag[tid] - no;
              L.icocc[slot] = mie + 1;
              iused_g[mie] = obase;
              for (int x = 0; x < no; ++x) {
                storej[obase + x] = L.cocc[loopj + obase + x];
                L.cocc[loopj + obase + x] = beta * L.cvir[loopi + mle + x];
                L.cvir[loopi + mle + x] = alpha * L.cvir[loopi + mle + x];
              }
            }
            const int grown_orbs = flag[D2_NT - 1];
            __syncthreads();
            if (s_fail) break;
            ncf_run += ngrow;
            mlf_run += grown_orbs;
            if (tid == 0) s_jext = max(s_jext, mlf_run);
          }
          mle_base += chunk_orbs;
          __syncthreads();
        }
        if (tid == 0) {
          s_ncf = ncf_run;
          s_mlf = mlf_run;
          s_mle = mle_base;
        }
        __syncthreads();
      }
      if (!s_fail) {
        if (tid == 0) L.ncf[j] = s_ncf;
        __syncthreads();
        // ---- occupied LMO atoms not in the virtual LMO: grow LMO i, in order
        const int nf = s_ncf;
        int mle_run = s_mle;
        int nce_run = s_nce;
        for (int c0 = 0; c0 < nf; c0 += D2_NT) {
          const int lf = c0 + tid;
          int ii = -1, no = 0, grow = 0;
          if (lf < nf) {
            ii = L.icocc[L.nncf[j] + lf] - 1;
            no = L.iorbs[ii];
            if (!latoms_g[ii]) {
              const int u = iused_g[ii];
              double s = 0.0;
              for (int x = 0; x < no; ++x) {
                const double t = beta * L.cocc[loopj + u + x];
                s += t * t;
              }
              grow = (s > P.thresh);
            }
          }
          flag[tid] = grow;
          __syncthreads();
          for (int off = 1; off < D2_NT; off <<= 1) {
            const int t = (tid >= off) ? flag[tid - off] : 0;
            __syncthreads();
            flag[tid] += t;
            __syncthreads();
          }
          const int ngrow = flag[D2_NT - 1];
          if (ngrow > 0) {
            const int rank = flag[tid] - grow;
            __syncthreads();
            flag[tid] = grow ? no : 0;
            __syncthreads();
            for (int off = 1; off < D2_NT; off <<= 1) {
              const int t = (tid >= off) ? flag[tid - off] : 0;
              __syncthreads();
              flag[tid] += t;
              __syncthreads();
            }
            if (grow) {
              const int slot = L.nnce[i] + nce_run + rank;
              const int obase = mle_run + flag[tid] - no;
              if (slot >= incv || ilr + obase + no > iur) atomicOr(&s_fail, 1);
            }
            __syncthreads();
            if (!s_fail && grow) {
              const int slot = L.nnce[i] + nce_run + rank;
              const int obase = mle_run + flag[tid] - no;
              const int u = iused_g[ii];
              L.icvir[slot] = ii + 1;
              latoms_g[ii] = 1;
              for (int x = 0; x < no; ++x) {
                storei[obase + x] = L.cvir[loopi + obase + x];
                L.cvir[loopi + obase + x] = -beta * L.cocc[loopj + u + x];
                L.cocc[loopj + u + x] = alpha * L.cocc[loopj + u + x];
              }
            }
            const int grown_orbs = flag[D2_NT - 1];
            __syncthreads();
            if (s_fail) break;
            nce_run += ngrow;
            mle_run += grown_orbs;
            if (tid == 0) s_iext = max(s_iext, mle_run);
          }
          __syncthreads();
        }
        if (!s_fail) {
          if (tid == 0) L.nce[i] = nce_run;
          __syncthreads();
          // reset the flags of the atoms now in the two LMOs
          for (int le = tid; le < L.nce[i]; le += D2_NT) latoms_g[L.icvir[L.nnce[i] + le] - 1] = 0;
          for (int lf = tid; lf < L.ncf[j]; lf += D2_NT) iused_g[L.icocc[L.nncf[j] + lf] - 1] = -1;
          __syncthreads();
          break;
        }
      }
      // ---- rejection (label 1000): restore both LMOs and all flags
      if (tid == 0) ++nrej;
      __syncthreads();
      const int rj = max(bj, s_jext), ri = max(bi, s_iext);
      for (int k = tid; k < rj; k += D2_NT) L.cocc[jlr + k] = storej[k];
      for (int k = tid; k < ri; k += D2_NT) L.cvir[ilr + k] = storei[k];
      for (int k = tid; k < L.numat; k += D2_NT) {
        iused_g[k] = -1;
        latoms_g[k] = 0;
      }
      __syncthreads();
      if (tid == 0) {
        L.ncf[j] = ncfj;
        L.nce[i] = ncei;
      }
      __syncthreads();
      if (P.retry) {
        alpha = 0.5 * (alpha + 1.0);
        continue;
      }
      break;
    }
    __syncthreads();
  }
  if (tid == 0) {
    sumb_out[ij] = sumb;
    nrej_out[ij] = nrej;
  }
}

// ---------------------------------------------------------------- diagg2, one workgroup
// The same rotations, every level of a diagg2 call in one launch: a single
// workgroup of D2W_NW waves on one CU, one wave per rotation, the waves
// taking the rotations of a level in turn and meeting at a workgroup barrier
// between levels.  All the waves share the CU's L1 and L2, so the barrier is
// the only synchronization, and a level costs one rotation's latency instead
// of a kernel launch.  Each wave keeps its atom maps in LDS (iused: offset of
// an atom's orbitals in occupied LMO j, 0xffff when absent; latoms: atom in
// virtual LMO i), entries set only for the atoms of the two LMOs and cleared
// after the rotation; scans and counts are wave shuffles and ballots.
constexpr int D2W_NW = 8;        // waves of the workgroup (at most)
constexpr unsigned short D2W_NONE = 0xffff;

__device__ __forceinline__ int wave_scan_incl(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// order of this wave's memory operations kept by the compiler across phases
// (the hardware keeps a wave's vector memory and LDS operations in order)
__device__ __forceinline__ void wave_phase() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// A rotation of the list with everything about it that no rotation of the
// call changes (its LMOs, their offsets and storage limits, the numerator
// and denominator of its angle), laid out by the host in level order so a
// rotation starts with one load.  Only the atom counts ncf[j] / nce[i], the
// lists and the coefficients are read as the earlier rotations left them.
struct D2Rot {
  double c, d;                  // fmo * damp, eigs[j] - eigv[i] - shift
  int ij, i, j;                 // list entry (0-based), virtual and occupied LMO
  int nnj, nni, jlr, ilr;       // atom and coefficient offsets of j and i
  int jur, iur, jncf, incv;     // their coefficient and atom limits (exclusive)
  int pad;
};

// Per-wave LDS of the rotation kernels: the first D2W_PCAP coefficients of
// each of the two LMOs as the rotation found them (cj, ci), the atom maps
// (iused, latoms) and copies of the two LMOs' atom lists (0-based atom
// numbers) and orbital counts when the two hold at most D2W_LCAP atoms
// together, which bounds both lists after the rotation (each ends as a
// subset of the union of the two); larger pairs read their lists from global
// memory.  Map entries are set for the atoms of the two LMOs when a rotation
// starts, so no clearing is needed.
constexpr int D2W_LCAP = 1024;
constexpr int D2W_PCAP = 1024;
struct D2Wave {
  double* cj;              // D2W_PCAP
  double* ci;              // D2W_PCAP
  unsigned short* iused;   // numat
  unsigned char* latoms;   // numat
  unsigned short* lj;      // D2W_LCAP
  unsigned short* li;      // D2W_LCAP
  unsigned char* oj;       // D2W_LCAP
  unsigned char* oi;       // D2W_LCAP
};

__host__ __device__ inline size_t d2w_wave_bytes(int numat) {
  return 16 * (size_t)D2W_PCAP + (((size_t)3 * numat + 15) & ~(size_t)15) + 6 * (size_t)D2W_LCAP;
}
// LDS bytes of a workgroup of nw waves
__host__ __device__ inline size_t d2w_lds_bytes(int numat, int nw) { return (size_t)nw * d2w_wave_bytes(numat); }

__device__ __forceinline__ D2Wave d2w_carve(unsigned char* lds, int numat, int w) {
  unsigned char* base = lds + (size_t)w * d2w_wave_bytes(numat);
  D2Wave v;
  v.cj = reinterpret_cast<double*>(base);
  v.ci = v.cj + D2W_PCAP;
  base += 16 * (size_t)D2W_PCAP;
  v.iused = reinterpret_cast<unsigned short*>(base);
  v.latoms = base + 2 * (size_t)numat;
  unsigned char* lists = base + (((size_t)3 * numat + 15) & ~(size_t)15);
  v.lj = reinterpret_cast<unsigned short*>(lists);
  v.li = v.lj + D2W_LCAP;
  v.oj = lists + 4 * (size_t)D2W_LCAP;
  v.oi = v.oj + D2W_LCAP;
  return v;
}

// dst[k] = src[k] for k < n by one wave, a pass of 64 * 8 elements at a
// time with all of a lane's loads issued before its stores and no branch
// between them (a lane past the end copies element n - 1 again: the same
// value to the same address), so a pass costs one memory round trip.
__device__ __forceinline__ void wave_copy(double* __restrict__ dst, const double* __restrict__ src, int n, int lane) {
  constexpr int U = 8;
  for (int k0 = 0; k0 < n; k0 += 64 * U) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = src[min(k0 + u * 64 + lane, n - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u) dst[min(k0 + u * 64 + lane, n - 1)] = v[u];
  }
}

// The rotation proper (after its lists are in place), on the LMOs' current
// coefficients read from the LDS copies (FAST: every position the rotation
// can touch lies inside them; each position is read before the rotation
// writes it, and a rejection restores the copies) or from global memory with
// an undo log (storej, storei).
template <bool FAST>
__device__ __forceinline__ void d2w_rotate(const LmoDev& L, const D2Rot& R, const D2Params& P, const D2Wave& M,
                                           double* __restrict__ storej, double* __restrict__ storei,
                                           double* __restrict__ sumb_out, int* __restrict__ nrej_out,
                                           const int lane, const int ncfj, const int ncei, const int bj,
                                           const int bi, const bool in_lds) {
  const unsigned long long below = (1ull << lane) - 1ull;
  double sumb = 0.0;
  int nrej = 0;
  const int i = R.i, j = R.j;
  const double c = R.c, d = R.d;
  const int nnj = R.nnj, nni = R.nni, jlr = R.jlr, ilr = R.ilr;
  const int jur = R.jur, iur = R.iur, jncf = R.jncf, incv = R.incv;
  const int* __restrict__ gj = L.icocc + nnj;
  const int* __restrict__ gi = L.icvir + nni;
  auto atom_j = [&](int k) { return in_lds ? (int)M.lj[k] : gj[k] - 1; };
  auto atom_i = [&](int k) { return in_lds ? (int)M.li[k] : gi[k] - 1; };
  auto orbs_j = [&](int k, int a) { return in_lds ? (int)M.oj[k] : L.iorbs[a]; };
  auto orbs_i = [&](int k, int a) { return in_lds ? (int)M.oi[k] : L.iorbs[a]; };
  // coefficient at position k of LMO j / i as the rotation found it
  auto old_j = [&](int k) -> double {
    if constexpr (FAST) return M.cj[k];
    else return L.cocc[jlr + k];
  };
  auto old_i = [&](int k) -> double {
    if constexpr (FAST) return M.ci[k];
    else return L.cvir[ilr + k];
  };
  const double e = copysign(sqrt(4.0 * c * c + d * d), d);
  double alpha = sqrt(0.5 * (1.0 + d / e));
  wave_phase();
  for (;;) {
    const double beta = -copysign(sqrt(1.0 - alpha * alpha), c);
    sumb += fabs(beta);
    // the maps for the atoms of the two LMOs: none of i's atoms in j yet,
    // none of j's in i
    for (int k = lane; k < ncei; k += 64) M.iused[atom_i(k)] = D2W_NONE;
    for (int k = lane; k < ncfj; k += 64) M.latoms[atom_j(k)] = 0;
    wave_phase();
    // offsets of LMO j's atoms
    int mlf = 0;
    for (int c0 = 0; c0 < ncfj; c0 += 64) {
      const int lf = c0 + lane;
      const int a = (lf < ncfj) ? atom_j(lf) : -1;
      const int no = (a >= 0) ? orbs_j(lf, a) : 0;
      const int incl = wave_scan_incl(no, lane);
      if (a >= 0) M.iused[a] = (unsigned short)(mlf + incl - no);
      mlf += __shfl(incl, 63, 64);
    }
    wave_phase();
    int jext = bj, iext = bi;
    bool fail = false;
    // virtual LMO atoms, in order: rotate common atoms, grow LMO j
    int mle_base = 0, mlf_run = mlf, ncf_run = ncfj;
    for (int c0 = 0; c0 < ncei; c0 += 64) {
      const int le = c0 + lane;
      int mie = -1, no = 0;
      if (le < ncei) {
        mie = atom_i(le);
        no = orbs_i(le, mie);
      }
      const int incl = wave_scan_incl(no, lane);
      const int mle = mle_base + incl - no;
      const int chunk_orbs = __shfl(incl, 63, 64);
      bool grow = false;
      double cv[4] = {0.0, 0.0, 0.0, 0.0};
      if (le < ncei) {
        M.latoms[mie] = 1;
        const int u = M.iused[mie];
        // an atom has 1 or 4 orbitals (checked when the LMO set is loaded).
        // Orbital x is read and written at min(x, no - 1), so the loads and
        // stores have no branch between them and every load of the lane is
        // issued before its stores; an s atom repeats orbital 0 (the same
        // value to the same address).
        const int xl = no - 1;
#pragma unroll
        for (int x = 0; x < 4; ++x) cv[x] = old_i(mle + min(x, xl));
        if (u != D2W_NONE) {
          double co[4];
#pragma unroll
          for (int x = 0; x < 4; ++x) co[x] = old_j(u + min(x, xl));
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            const double a = co[x], b = cv[x];
            L.cocc[jlr + u + min(x, xl)] = alpha * a + beta * b;
            L.cvir[ilr + mle + min(x, xl)] = alpha * b - beta * a;
          }
        } else {
          double s2 = 0.0;
          for (int x = 0; x < no; ++x) {
            const double t = beta * cv[x];
            s2 += t * t;
          }
          grow = (s2 > P.thresh);
        }
      }
      const unsigned long long gm = __ballot(grow);
      if (gm) {
        const int rank = __popcll(gm & below);
        const int gincl = wave_scan_incl(grow ? no : 0, lane);
        const int slot = nnj + ncf_run + rank;
        const int obase = mlf_run + gincl - no;
        // the reference stops at the first atom that would not fit
        fail = __any(grow && (slot >= jncf || jlr + obase + no > jur));
        if (!fail && grow) {
          L.icocc[slot] = mie + 1;
          if (in_lds) {
            M.lj[ncf_run + rank] = (unsigned short)mie;
            M.oj[ncf_run + rank] = (unsigned char)no;
          }
          M.iused[mie] = (unsigned short)obase;
          const int xl = no - 1;
          if constexpr (!FAST) {
            double old[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) old[x] = L.cocc[jlr + obase + min(x, xl)];
#pragma unroll
            for (int x = 0; x < 4; ++x) storej[obase + min(x, xl)] = old[x];
          }
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            L.cocc[jlr + obase + min(x, xl)] = beta * cv[x];
            L.cvir[ilr + mle + min(x, xl)] = alpha * cv[x];
          }
        }
        if (fail) break;
        ncf_run += __popcll(gm);
        mlf_run += __shfl(gincl, 63, 64);
        jext = max(jext, mlf_run);
      }
      mle_base += chunk_orbs;
    }
    if (in_lds) wave_phase();
    else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // grown atoms re-read from global
    if (!fail) {
      if (lane == 0) L.ncf[j] = ncf_run;
      // occupied LMO atoms not in the virtual LMO: grow LMO i, in order
      int mle_run = mle_base, nce_run = ncei;
      for (int c0 = 0; c0 < ncf_run; c0 += 64) {
        const int lf = c0 + lane;
        int ii = -1, no = 0, u = 0;
        bool grow = false;
        double cj[4] = {0.0, 0.0, 0.0, 0.0};
        if (lf < ncf_run) {
          ii = atom_j(lf);
          no = orbs_j(lf, ii);
          if (!M.latoms[ii]) {
            // an atom of LMO j as the rotation found it (not touched above)
            u = M.iused[ii];
#pragma unroll
            for (int x = 0; x < 4; ++x) cj[x] = old_j(u + min(x, no - 1));
            double s2 = 0.0;
            for (int x = 0; x < no; ++x) {
              const double t = beta * cj[x];
              s2 += t * t;
            }
            grow = (s2 > P.thresh);
          }
        }
        const unsigned long long gm = __ballot(grow);
        if (gm) {
          const int rank = __popcll(gm & below);
          const int gincl = wave_scan_incl(grow ? no : 0, lane);
          const int slot = nni + nce_run + rank;
          const int obase = mle_run + gincl - no;
          fail = __any(grow && (slot >= incv || ilr + obase + no > iur));
          if (!fail && grow) {
            L.icvir[slot] = ii + 1;
            if (in_lds) M.li[nce_run + rank] = (unsigned short)ii;
            M.latoms[ii] = 1;
            const int xl = no - 1;
            if constexpr (!FAST) {
              double old[4];
#pragma unroll
              for (int x = 0; x < 4; ++x) old[x] = L.cvir[ilr + obase + min(x, xl)];
#pragma unroll
              for (int x = 0; x < 4; ++x) storei[obase + min(x, xl)] = old[x];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x) {
              L.cvir[ilr + obase + min(x, xl)] = -beta * cj[x];
              L.cocc[jlr + u + min(x, xl)] = alpha * cj[x];
            }
          }
          if (fail) break;
          nce_run += __popcll(gm);
          mle_run += __shfl(gincl, 63, 64);
          iext = max(iext, mle_run);
        }
      }
      if (!fail) {
        if (lane == 0) L.nce[i] = nce_run;
        wave_phase();
        break;
      }
    }
    // rejection (label 1000): restore both LMOs as the rotation found them
    // (the undo log is written by other lanes: wait for the wave's stores
    // first); the maps are set again at the retry
    ++nrej;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const int rj = max(bj, jext), ri = max(bi, iext);
    if constexpr (FAST) {
      for (int k = lane; k < rj; k += 64) L.cocc[jlr + k] = M.cj[k];
      for (int k = lane; k < ri; k += 64) L.cvir[ilr + k] = M.ci[k];
    } else {
      for (int k = lane; k < rj; k += 64) L.cocc[jlr + k] = storej[k];
      for (int k = lane; k < ri; k += 64) L.cvir[ilr + k] = storei[k];
    }
    if (lane == 0) {
      L.ncf[j] = ncfj;
      L.nce[i] = ncei;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (P.retry) {
      alpha = 0.5 * (alpha + 1.0);
      continue;
    }
    break;
  }
  if (lane == 0) {
    sumb_out[R.ij] = sumb;
    nrej_out[R.ij] = nrej;
  }
}

// One rotation by one wave.  The dependent global reads are the record, then
// in one round trip the two atom counts together with the heads of the two
// atom lists and the LMOs' coefficients (read speculatively, clamped to each
// LMO's storage), then the orbital counts of the listed atoms; when the
// coefficient copies in LDS cover every position the rotation can touch
// (FAST), the rotation itself reads nothing more from global memory.
__device__ void d2w_rotation(const LmoDev& L, const D2Rot& R, const D2Params& P, const D2Wave& M,
                             double* __restrict__ storej, double* __restrict__ storei, double* __restrict__ sumb_out,
                             int* __restrict__ nrej_out, const int lane) {
  const int i = R.i, j = R.j;
  const int nnj = R.nnj, nni = R.nni, jlr = R.jlr, ilr = R.ilr;
  const int jur = R.jur, iur = R.iur, jncf = R.jncf, incv = R.incv;
  const int* __restrict__ gj = L.icocc + nnj;
  const int* __restrict__ gi = L.icvir + nni;
  const int ncap_j = jncf - nnj, ncap_i = incv - nni;
  const int pj = min(jur - jlr, P.pcap), pi = min(iur - ilr, P.pcap);
  const bool spec = ncap_j > 0 && ncap_i > 0 && pj > 0 && pi > 0;
  int ncfj, ncei, sj = 0, si = 0;
  if (spec) {
    constexpr int U = D2W_PCAP / 64;
    double uj[U], ui[U];
    ncfj = L.ncf[j];
    ncei = L.nce[i];
    sj = gj[min(lane, ncap_j - 1)];
    si = gi[min(lane, ncap_i - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uj[u] = L.cocc[jlr + min(u * 64 + lane, pj - 1)];
      ui[u] = L.cvir[ilr + min(u * 64 + lane, pi - 1)];
    }
    // a lane past the end copies the last position again (the same value)
#pragma unroll
    for (int u = 0; u < U; ++u) {
      M.cj[min(u * 64 + lane, pj - 1)] = uj[u];
      M.ci[min(u * 64 + lane, pi - 1)] = ui[u];
    }
  } else {
    ncfj = L.ncf[j];
    ncei = L.nce[i];
  }
  const int bj = min(4 * ncfj, jur - jlr), bi = min(4 * ncei, iur - ilr);
  // atom lists (and orbital counts) into LDS when they fit (every list entry
  // the rotation reads lies below ncfj + ncei)
  const bool in_lds = ncfj + ncei <= P.lcap;
  // every position the rotation can touch: below the LMO's storage and below
  // 4 orbitals for each atom of the two LMOs
  const int reach = 4 * (ncfj + ncei);
  const bool fast = spec && in_lds && min(jur - jlr, reach) <= pj && min(iur - ilr, reach) <= pi;
  if (!fast) {
    // undo log of the coefficients the LMOs hold now
    wave_copy(storej, L.cocc + jlr, bj, lane);
    wave_copy(storei, L.cvir + ilr, bi, lane);
  }
  if (in_lds && ncfj > 0 && ncei > 0) {
    // the first 64 entries from the speculative reads, the rest read now,
    // branch-free (a lane past the end of a list repeats its last entry:
    // the same value to the same place)
    int kj = min(lane, ncfj - 1), ki = min(lane, ncei - 1);
    int aj = (spec ? __shfl(sj, kj, 64) : gj[kj]) - 1;
    int ai = (spec ? __shfl(si, ki, 64) : gi[ki]) - 1;
    for (int k0 = 0;;) {
      const int nj = L.iorbs[aj], ni = L.iorbs[ai];
      M.lj[kj] = (unsigned short)aj;
      M.oj[kj] = (unsigned char)nj;
      M.li[ki] = (unsigned short)ai;
      M.oi[ki] = (unsigned char)ni;
      k0 += 64;
      if (k0 >= max(ncfj, ncei)) break;
      kj = min(k0 + lane, ncfj - 1);
      ki = min(k0 + lane, ncei - 1);
      aj = gj[kj] - 1;
      ai = gi[ki] - 1;
    }
  } else if (in_lds) {
    for (int k = lane; k < ncfj; k += 64) M.lj[k] = (unsigned short)(gj[k] - 1);
    for (int k = lane; k < ncei; k += 64) M.li[k] = (unsigned short)(gi[k] - 1);
    for (int k = lane; k < ncfj; k += 64) M.oj[k] = (unsigned char)L.iorbs[M.lj[k]];
    for (int k = lane; k < ncei; k += 64) M.oi[k] = (unsigned char)L.iorbs[M.li[k]];
  }
  if (fast)
    d2w_rotate<true>(L, R, P, M, storej, storei, sumb_out, nrej_out, lane, ncfj, ncei, bj, bi, in_lds);
  else
    d2w_rotate<false>(L, R, P, M, storej, storei, sumb_out, nrej_out, lane, ncfj, ncei, bj, bi, in_lds);
}

// Every level of a run in one workgroup: the waves take a level's rotations
// in turn and meet at the barrier before the next level.
__global__ __launch_bounds__(64 * D2W_NW) void k_mz_diagg2_wg(LmoDev L, const D2Rot* __restrict__ rec, D2Params P,
                                                            const int* __restrict__ level_start, int lev0,
                                                            int lev1, double* __restrict__ storej_all, int span_j,
                                                            double* __restrict__ storei_all, int span_i,
                                                            double* __restrict__ sumb_out,
                                                            int* __restrict__ nrej_out) {
  extern __shared__ __align__(16) unsigned char d2w_lds[];
  const int nw = blockDim.x >> 6;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const D2Wave M = d2w_carve(d2w_lds, L.numat, w);
  double* storej = storej_all + (size_t)w * span_j;
  double* storei = storei_all + (size_t)w * span_i;
  for (int l = lev0; l < lev1; ++l) {
    const int b1 = level_start[l + 1];
    for (int b = level_start[l] + w; b < b1; b += nw) d2w_rotation(L, rec[b], P, M, storej, storei, sumb_out, nrej_out, lane);
    __syncthreads();  // every wave reaches it once per level
  }
}

// One level across the chip: a one-wave workgroup per rotation.
__global__ __launch_bounds__(64) void k_mz_diagg2_lvl(LmoDev L, const D2Rot* __restrict__ rec, D2Params P,
                                                      int base, double* __restrict__ storej_all, int span_j,
                                                      double* __restrict__ storei_all, int span_i,
                                                      double* __restrict__ sumb_out, int* __restrict__ nrej_out) {
  extern __shared__ __align__(16) unsigned char d2w_lds[];
  const D2Wave M = d2w_carve(d2w_lds, L.numat, 0);
  const int slot = blockIdx.x;
  d2w_rotation(L, rec[base + slot], P, M, storej_all + (size_t)slot * span_j, storei_all + (size_t)slot * span_i,
               sumb_out, nrej_out, threadIdx.x);
}

// ---------------------------------------------------------------- helecz
// One workgroup per atom i: sum over its stored pairs (j < i) and its
// diagonal block; partial sums combined in a fixed order by k_mz_sum2.
__global__ __launch_bounds__(256) void k_mz_helecz(int numat, const int* __restrict__ iorbs,
                                                   const int* __restrict__ nijbo,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ h,
                                                   const double* __restrict__ f,
                                                   double* __restrict__ part) {
  const int i = blockIdx.x, tid = threadIdx.x;
  const int ni = iorbs[i];
  double ee = 0.0, ed = 0.0;
  for (int j = tid; j < i; j += 256) {
    const int k = nijbo[(size_t)i + (size_t)j * numat];
    if (k < 0) continue;
    const int len = ni * iorbs[j];
    for (int m = k; m < k + len; ++m) ee += p[m] * (h[m] + f[m]);
  }
  if (tid == 0) {
    int k = nijbo[(size_t)i + (size_t)i * numat];
    for (int l1 = 0; l1 < ni; ++l1) {
      for (int l2 = 0; l2 < l1; ++l2, ++k) ee += p[k] * (h[k] + f[k]);
      ed += p[k] * (h[k] + f[k]);
      ++k;
    }
  }
  __shared__ double r1[256], r2[256];
  r1[tid] = ee;
  r2[tid] = ed;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      r1[tid] += r1[tid + w];
      r2[tid] += r2[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[2 * i] = r1[0];
    part[2 * i + 1] = r2[0];
  }
}

__global__ __launch_bounds__(1024) void k_mz_sum2(int n, const double* __restrict__ part,
                                                  double* __restrict__ out) {
  __shared__ double r1[1024], r2[1024];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  r1[threadIdx.x] = a;
  r2[threadIdx.x] = b;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = r1[0] + 0.5 * r2[0];
}

// ---------------------------------------------------------------- cnvgz
// pmax partials over the packed vectors
__global__ __launch_bounds__(256) void k_mz_absdiff_max(long long n, const double* __restrict__ a,
                                                        const double* __restrict__ b,
                                                        double* __restrict__ part) {
  __shared__ double r[256];
  double m = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    m = fmax(m, fabs(a[i] - b[i]));
  r[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) r[threadIdx.x] = fmax(r[threadIdx.x], r[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

// One workgroup: diagonal bookkeeping, pmax, and the extrapolation factor
// (cnvgz.F90:25-54); scal[0] = pmax, scal[1] = fac (0 when not applied).
__global__ __launch_bounds__(1024) void k_mz_cnvgz_diag(int norbs, int nparts, const double* __restrict__ part,
                                                        const int* __restrict__ idiag,
                                                        const double* __restrict__ pnew,
                                                        const double* __restrict__ p,
                                                        const double* __restrict__ p1, double* __restrict__ p2,
                                                        double* __restrict__ p3, int niter, int three,
                                                        double* __restrict__ scal) {
  __shared__ double ra[1024], rb[1024];
  const int tid = threadIdx.x;
  double fa = 0.0, fb = 0.0, m = 0.0;
  for (int i = tid; i < norbs; i += 1024) {
    const int j = idiag[i] - 1;
    const double a3 = pnew[j], a2 = p[j];
    p3[i] = a3;
    p2[i] = a2;
    const double sa = fabs(a3 - a2);
    fa += sa * sa;
    const double t = a3 - 2.0 * a2 + p1[i];
    fb += t * t;
  }
  for (int k = tid; k < nparts; k += 1024) m = fmax(m, part[k]);
  ra[tid] = fa;
  rb[tid] = fb;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (tid < w) {
      ra[tid] += ra[tid + w];
      rb[tid] += rb[tid + w];
    }
    __syncthreads();
  }
  const double faca = ra[0], facb = rb[0];
  __syncthreads();
  ra[tid] = m;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (tid < w) ra[tid] = fmax(ra[tid], ra[tid + w]);
    __syncthreads();
  }
  if (tid == 0) {
    scal[0] = ra[0];
    double fac = 0.0;
    if (three && niter % 3 == 0 && facb > 0.0 && faca < 100.0 * facb) fac = sqrt(faca / facb);
    scal[1] = fac;
  }
}

// pnew += fac (pnew - p); then p = pnew (after the damping pass)
__global__ void k_mz_extrap(long long n, const double* __restrict__ scal, double* __restrict__ pnew,
                            const double* __restrict__ p) {
  const double fac = scal[1];
  if (fac == 0.0) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    pnew[i] = pnew[i] + fac * (pnew[i] - p[i]);
}

// diagonal damping from iteration 4 on (cnvgz.F90:60-74), then p1 = p2
__global__ void k_mz_damp(int norbs, const int* __restrict__ idiag, const double* __restrict__ scal,
                          const double* __restrict__ p2, const double* __restrict__ p3, int niter, int three,
                          double* __restrict__ pnew, double* __restrict__ p1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= norbs) return;
  const double damp = 0.05;
  if (three && niter > 3 && scal[0] > damp) {
    const int j = idiag[i] - 1;
    const double d = p3[i] - p2[i];
    if (fabs(d) > damp) {
      double v = p2[i] + copysign(damp, d);
      pnew[j] = fmin(2.0, fmax(v, 0.0));
    }
  }
  p1[i] = p2[i];
}


// ---------------------------------------------------------------- tidy, check, eimp
// tidy (src/MOZYME/tidy.F90:16-244), first pass, one thread per LMO in the
// reference's order: an atom stays when the squared sum of its coefficients
// exceeds thresh; its flag, and the LMO's kept atom and coefficient counts.
__global__ void k_mz_tidy_count(int nmo, const int* __restrict__ nc, const int* __restrict__ nnc,
                                const int* __restrict__ ncmo, const int* __restrict__ ic,
                                const double* __restrict__ c, const int* __restrict__ iorbs, double thresh,
                                unsigned char* __restrict__ keep, int* __restrict__ kc, int* __restrict__ kq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmo) return;
  int m = ncmo[i], kept = 0, q = 0;
  for (int k = 0; k < nc[i]; ++k) {
    const int no = iorbs[ic[nnc[i] + k] - 1];
    double sum = 0.0;
    for (int x = 0; x < no; ++x) sum += c[m + x] * c[m + x];
    const bool kk = sum > thresh;
    keep[nnc[i] + k] = kk ? 1 : 0;
    if (kk) {
      ++kept;
      q += no;
    }
    m += no;
  }
  kc[i] = kept;
  kq[i] = q;
}

// tidy's moves: the kept atoms and coefficients of each LMO, in order, from
// its old place in (ic, c) to its new one in (ic2, c2); the count becomes
// the kept count.  Only each LMO's own extent is written.
__global__ void k_mz_tidy_move(int nmo, int* __restrict__ nc, const int* __restrict__ nnc,
                               const int* __restrict__ ncmo, const int* __restrict__ ic,
                               const double* __restrict__ c, const int* __restrict__ iorbs,
                               const unsigned char* __restrict__ keep, const int* __restrict__ nnc_new,
                               const int* __restrict__ ncmo_new, int* __restrict__ ic2, double* __restrict__ c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmo) return;
  int m = ncmo[i], r = 0, q = ncmo_new[i];
  for (int k = 0; k < nc[i]; ++k) {
    const int a = ic[nnc[i] + k];
    const int no = iorbs[a - 1];
    if (keep[nnc[i] + k]) {
      ic2[nnc_new[i] + r++] = a;
      for (int x = 0; x < no; ++x) c2[q + x] = c[m + x];
      q += no;
    }
    m += no;
  }
  nc[i] = r;
}

// check (src/MOZYME/check.F90:16-113): each LMO's squared norm, summed in the
// reference's order by one thread, |1 - norm| into err, then the LMO scaled
// to unit norm.
__global__ void k_mz_check(int nmo, const int* __restrict__ nc, const int* __restrict__ nnc,
                           const int* __restrict__ ncmo, const int* __restrict__ ic, double* __restrict__ c,
                           const int* __restrict__ iorbs, double* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmo) return;
  int n = 0;
  double sum = 0.0;
  for (int k = 0; k < nc[i]; ++k) {
    const int no = iorbs[ic[nnc[i] + k] - 1];
    for (int x = 0; x < no; ++x, ++n) sum += c[ncmo[i] + n] * c[ncmo[i] + n];
  }
  err[i] = fabs(1.0 - sum);
  const double sc = 1.0 / sqrt(sum);
  for (int x = 0; x < n; ++x) c[ncmo[i] + x] *= sc;
}

// eimp (src/MOZYME/eimp.F90:16-45): the first element of each stored
// off-diagonal block of p becomes the sum of squares of that block of f.
__global__ void k_mz_eimp(int nblk, const int* __restrict__ off, const int* __restrict__ len,
                          const double* __restrict__ f, double* __restrict__ p) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  double sum = 0.0;
  for (int m = 0; m < len[b]; ++m) sum += f[off[b] + m] * f[off[b] + m];
  p[off[b]] = sum;
}

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace
}  // namespace mhip

using namespace mhip;

// ---- host side: device copies of an LMO set ---------------------------------
namespace {

struct LmoBufs {
  DevBuf nfirst, iorbs, nijbo, ncf, nncf, ncocc, icocc, cocc, nce, nnce, ncvir, icvir, cvir;
  LmoDev dev{};
};

// Optional HIP-event timing and counts of the diagg2 levels (process-wide).
struct MzStats {
  bool on = false;
  double ms = 0.0, bytes = 0.0;
  long long rotations = 0, levels = 0, launches = 0, rounds = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};
MzStats g_mz_stats;
// diagg2 on the LDS rotation kernels (the default: levels of more than one
// round of the workgroup across the chip, k_mz_diagg2_lvl; runs of narrower
// levels in one workgroup, k_mz_diagg2_wg) or on the global-map kernel one
// launch per level (k_mz_diagg2).  Measured on the 400-residue helix in the
// reference order: 0.86 s of rotations per SCF against 1.18 s.
bool g_mz_one_wg = true;
int g_mz_wide_rounds = 1;
int g_mz_lcap = 1024;  // D2W_LCAP
int g_mz_pcap = 1024;  // D2W_PCAP

// Device copies and scratch of the MOZYME entry points, one set per device,
// grown as needed and kept until the process ends: no allocation per call.
// Grow-only pinned host memory, kept for the life of the process.
struct HostPin {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* ensure(size_t need) {
    if (need > bytes) {
      if (ptr) MHIP_CHECK(hipHostFree(ptr));
      ptr = nullptr;
      bytes = 0;
      MHIP_CHECK(hipHostMalloc(&ptr, need));
      bytes = need;
    }
    return ptr;
  }
};

struct MzWork {
  LmoBufs B;
  HostPin hrec, hsumb, hnrej;  // diagg2's rotation records and per-rotation results
  std::vector<int> nijbo_seen, iorbs_seen;  // the nijbo (and sizes) last checked and uploaded
  long long nijbo_mpack = -1;
  bool resident = false;                 // a resident MOZYME loop owns B
  std::vector<int> nq_occ, nq_vir;       // its LMOs' coefficient counts (last tidy)
  DevBuf icocc2, cocc2, icvir2, cvir2;   // tidy's second copies (swapped with B's)
  DevBuf dkeep, dkc, dkq, dnn_new, dnq_new, derr, deoff, delen;
  DevBuf dp, dpart, dfao, daocc, dcounts, doffs, dnfmo, dold, dfmo, difmo, dvmax, dvsum, deigv, deigs, dorder,
      dused, dlat, dsj, dsi, dsumb, dnrej, dh, df, dout, dpn, d1, d2, d3, di, dsc, dlevel, drec;
  // the stored atom pairs (a >= b) of nijbo, and the density gather's lists
  DevBuf dspairs, dacnt, daoff, dalmo, dacoff, dbits, dcum;
  int nspairs = 0;
  // diagg1's candidates: stored neighbours per atom (fixed per nijbo) and the
  // occupied LMOs by first and second atom (per call)
  DevBuf dnbr_off, dnbr, dfl_cnt, dfl_off, dfl_ent;
};
MzWork& mz_work() {
  static std::map<int, MzWork> work;
  int d = 0;
  MHIP_CHECK(hipGetDevice(&d));
  return work[d];
}

void up(hipStream_t s, DevBuf& b, const void* src, size_t bytes) {
  b.ensure(bytes ? bytes : 8);
  if (bytes) MHIP_CHECK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s));
}
void down(hipStream_t s, void* dst, const DevBuf& b, size_t bytes) {
  if (bytes) MHIP_CHECK(hipMemcpyAsync(dst, b.ptr, bytes, hipMemcpyDeviceToHost, s));
}

// nijbo into W.B.nijbo (and iorbs into W.B.iorbs), checked first: every
// block inside mpack, the matrix symmetric, every diagonal block stored.
// A nijbo identical to the last one checked is neither checked nor sent again.
void upload_nijbo(hipStream_t s, MzWork& W, int numat, const int* nijbo, long long mpack,
                  const std::vector<int>& iorbs) {
  const size_t nn = (size_t)numat * numat;
  if (W.nijbo_seen.size() == nn && W.nijbo_mpack == mpack && W.iorbs_seen == iorbs &&
      std::memcmp(W.nijbo_seen.data(), nijbo, sizeof(int) * nn) == 0)
    return;
  for (int b = 0; b < numat; ++b)
    for (int a = 0; a < numat; ++a) {
      const long long o = nijbo[(size_t)a + (size_t)b * numat];
      if (o < 0) {
        MHIP_REQUIRE(a != b, "LMO set: every diagonal block of nijbo must be stored");
        continue;
      }
      const long long blk = (a == b) ? iorbs[a] * (iorbs[a] + 1) / 2 : (long long)iorbs[a] * iorbs[b];
      MHIP_REQUIRE(o + blk <= mpack, "LMO set: nijbo block beyond mpack");
      MHIP_REQUIRE(nijbo[(size_t)b + (size_t)a * numat] == o, "LMO set: nijbo not symmetric");
    }
  W.nijbo_seen.clear();  // stays empty if an upload below fails
  std::vector<int> sp;
  for (int b = 0; b < numat; ++b)
    for (int a = b; a < numat; ++a)
      if (nijbo[(size_t)a + (size_t)b * numat] >= 0) {
        sp.push_back(a);
        sp.push_back(b);
      }
  W.nspairs = (int)(sp.size() / 2);
  up(s, W.dspairs, sp.data(), sizeof(int) * sp.size());
  std::vector<int> noff(numat + 1, 0), nb;
  for (int a = 0; a < numat; ++a) {
    noff[a] = (int)nb.size();
    for (int b = 0; b < numat; ++b)
      if (nijbo[(size_t)a + (size_t)b * numat] >= 0) nb.push_back(b);
  }
  noff[numat] = (int)nb.size();
  up(s, W.dnbr_off, noff.data(), sizeof(int) * noff.size());
  up(s, W.dnbr, nb.data(), sizeof(int) * std::max<size_t>(nb.size(), 1));
  up(s, W.B.nijbo, nijbo, sizeof(int) * nn);
  up(s, W.B.iorbs, iorbs.data(), sizeof(int) * numat);
  MHIP_CHECK(hipStreamSynchronize(s));
  W.nijbo_seen.assign(nijbo, nijbo + nn);
  W.iorbs_seen = iorbs;
  W.nijbo_mpack = mpack;
}

// Checks every index the kernels will follow, then uploads.
void upload_lmos(hipStream_t s, const mhip_lmo_set* L, MzWork& W) {
  LmoBufs& B = W.B;
  MHIP_REQUIRE(L && L->numat > 0 && L->norbs > 0 && L->nocc >= 0 && L->nvir >= 0, "LMO set: bad sizes");
  std::vector<int> iorbs(L->numat);
  for (int a = 0; a < L->numat; ++a) {
    iorbs[a] = L->nlast[a] - L->nfirst[a] + 1;
    MHIP_REQUIRE(iorbs[a] == 1 || iorbs[a] == 4, "LMO set: only s and sp atoms");
  }
  auto check_set = [&](int n, const int* nc, const int* nn, const int* no, const int* idx, int idim,
                       int cdim, const char* what) {
    for (int i = 0; i < n; ++i) {
      const int end_atoms = (i + 1 < n) ? nn[i + 1] : idim;
      const int end_coef = (i + 1 < n) ? no[i + 1] : cdim;
      MHIP_REQUIRE(nn[i] >= 0 && nc[i] >= 0 && nn[i] + nc[i] <= end_atoms && end_atoms <= idim,
                   std::string("LMO set: atom list out of range in ") + what);
      MHIP_REQUIRE(nc[i] <= MZ_MAXAT, std::string("LMO set: more than 256 atoms in one LMO in ") + what);
      int orbs = 0;
      for (int k = 0; k < nc[i]; ++k) {
        const int a = idx[nn[i] + k];
        MHIP_REQUIRE(a >= 1 && a <= L->numat, std::string("LMO set: bad atom number in ") + what);
        orbs += iorbs[a - 1];
      }
      MHIP_REQUIRE(no[i] >= 0 && no[i] + orbs <= end_coef && end_coef <= cdim,
                   std::string("LMO set: coefficients out of range in ") + what);
    }
  };
  check_set(L->nocc, L->ncf, L->nncf, L->ncocc, L->icocc, L->icocc_dim, L->cocc_dim, "the occupied set");
  check_set(L->nvir, L->nce, L->nnce, L->ncvir, L->icvir, L->icvir_dim, L->cvir_dim, "the virtual set");
  upload_nijbo(s, W, L->numat, L->nijbo, L->mpack, iorbs);
  up(s, B.nfirst, L->nfirst, sizeof(int) * L->numat);
  up(s, B.ncf, L->ncf, sizeof(int) * L->nocc);
  up(s, B.nncf, L->nncf, sizeof(int) * L->nocc);
  up(s, B.ncocc, L->ncocc, sizeof(int) * L->nocc);
  up(s, B.icocc, L->icocc, sizeof(int) * L->icocc_dim);
  up(s, B.cocc, L->cocc, sizeof(double) * L->cocc_dim);
  up(s, B.nce, L->nce, sizeof(int) * L->nvir);
  up(s, B.nnce, L->nnce, sizeof(int) * L->nvir);
  up(s, B.ncvir, L->ncvir, sizeof(int) * L->nvir);
  up(s, B.icvir, L->icvir, sizeof(int) * L->icvir_dim);
  up(s, B.cvir, L->cvir, sizeof(double) * L->cvir_dim);
  LmoDev& d = B.dev;
  d.numat = L->numat;
  d.norbs = L->norbs;
  d.nocc = L->nocc;
  d.nvir = L->nvir;
  d.nfirst = B.nfirst.as<int>();
  d.iorbs = B.iorbs.as<int>();
  d.nijbo = B.nijbo.as<int>();
  d.ncf = B.ncf.as<int>();
  d.nncf = B.nncf.as<int>();
  d.ncocc = B.ncocc.as<int>();
  d.icocc = B.icocc.as<int>();
  d.icocc_dim = L->icocc_dim;
  d.cocc = B.cocc.as<double>();
  d.cocc_dim = L->cocc_dim;
  d.nce = B.nce.as<int>();
  d.nnce = B.nnce.as<int>();
  d.ncvir = B.ncvir.as<int>();
  d.icvir = B.icvir.as<int>();
  d.icvir_dim = L->icvir_dim;
  d.cvir = B.cvir.as<double>();
  d.cvir_dim = L->cvir_dim;
}

template <class F>
int mz_guard(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    return 1;
  }
}

hipStream_t g_mz_stream = nullptr;
hipStream_t mz_stream(mhip_ctx* ctx) {
  (void)ctx;
  if (!g_mz_stream) MHIP_CHECK(hipStreamCreateWithFlags(&g_mz_stream, hipStreamNonBlocking));
  return g_mz_stream;
}

}  // namespace


// ---- device-pointer cores of the entry points (the resident MOZYME loop
// calls them directly; the entry points add the host transfers) --------------
namespace mhip {

void mz_density_dev(hipStream_t s, long long mpack, int nocc, int mode, const double* partp_d, double* p_d) {
  MzWork& W = mz_work();
  const LmoBufs& B = W.B;
  const int nb = std::min(4096, cdiv(mpack, 256));
  const double spinfa = (mode == 0 || mode == 1) ? 2.0 : (mode == -1 ? -2.0 : 1.0);
  const bool scale = std::fabs(spinfa - 1.0) > 1e-10;
  // every element starts from its mode's value (elements of no stored pair
  // keep it, scaled like the rest)
  hipLaunchKernelGGL(k_mz_density_init, dim3(nb), dim3(256), 0, s, mpack, mode, partp_d, p_d);
  MHIP_LAUNCH_CHECK();
  if (scale) {
    hipLaunchKernelGGL(k_scale, dim3(nb), dim3(256), 0, s, mpack, spinfa, p_d);
    MHIP_LAUNCH_CHECK();
  }
  if (nocc <= 0 || W.nspairs == 0) return;
  const int numat = B.dev.numat;
  const int nw = cdiv(nocc, 64);
  const long long nbits = (long long)numat * nw;
  W.dbits.ensure(sizeof(unsigned long long) * nbits);
  W.dcum.ensure(sizeof(int) * nbits);
  W.dacnt.ensure(sizeof(int) * (numat + 1));
  W.daoff.ensure(sizeof(int) * (numat + 1));
  W.dalmo.ensure(sizeof(int) * std::max(B.dev.icocc_dim, 1));
  W.dacoff.ensure(sizeof(int) * std::max(B.dev.icocc_dim, 1));
  auto* bits = W.dbits.as<unsigned long long>();
  hipLaunchKernelGGL(k_mz_fill_u64, dim3((int)std::min<long long>(4096, cdiv(nbits, 256))), dim3(256), 0, s, nbits,
                     bits);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_atom_bits, dim3(cdiv(nocc, 256)), dim3(256), 0, s, B.dev, nw, bits);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_atom_cum, dim3(cdiv(numat, 256)), dim3(256), 0, s, numat, nw, bits, W.dcum.as<int>(),
                     W.dacnt.as<int>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_scan_counts, dim3(1), dim3(1024), 0, s, numat, W.dacnt.as<int>(), W.daoff.as<int>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_atom_fill, dim3(cdiv(nocc, 256)), dim3(256), 0, s, B.dev, nw, bits, W.dcum.as<int>(),
                     W.daoff.as<int>(), W.dalmo.as<int>(), W.dacoff.as<int>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_density_gather, dim3(cdiv(W.nspairs, 128)), dim3(128), 0, s, B.dev, W.nspairs,
                     W.dspairs.as<int2>(), W.daoff.as<int>(), W.dalmo.as<int>(), W.dacoff.as<int>(), mode, partp_d,
                     spinfa, scale ? 1 : 0, p_d);
  MHIP_LAUNCH_CHECK();
}

double mz_helecz_dev(hipStream_t s, int numat, const double* p_d, const double* h_d, const double* f_d) {
  MzWork& W = mz_work();
  W.dpart.ensure(sizeof(double) * 2 * numat);
  W.dout.ensure(sizeof(double));
  hipLaunchKernelGGL(k_mz_helecz, dim3(numat), dim3(256), 0, s, numat, W.B.iorbs.as<int>(), W.B.nijbo.as<int>(),
                     p_d, h_d, f_d, W.dpart.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_sum2, dim3(1), dim3(1024), 0, s, numat, W.dpart.as<double>(), W.dout.as<double>());
  MHIP_LAUNCH_CHECK();
  double e = 0.0;
  down(s, &e, W.dout, sizeof(double));
  MHIP_CHECK(hipStreamSynchronize(s));
  return e;
}

// pnew and p (device) end equal, as cnvgz leaves them
double mz_cnvgz_dev(hipStream_t s, int norbs, long long mpack, double* pn_d, double* p_d, double* p1_d,
                    double* p2_d, double* p3_d, int niter, const int* idiag_d, int three_point) {
  MzWork& W = mz_work();
  const int nparts = std::min(1024, cdiv(mpack, 256));
  W.dpart.ensure(sizeof(double) * nparts);
  W.dsc.ensure(sizeof(double) * 2);
  hipLaunchKernelGGL(k_mz_absdiff_max, dim3(nparts), dim3(256), 0, s, mpack, pn_d, p_d, W.dpart.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_cnvgz_diag, dim3(1), dim3(1024), 0, s, norbs, nparts, W.dpart.as<double>(), idiag_d, pn_d,
                     p_d, p1_d, p2_d, p3_d, niter, three_point, W.dsc.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_extrap, dim3(std::min(4096, cdiv(mpack, 256))), dim3(256), 0, s, mpack,
                     W.dsc.as<double>(), pn_d, p_d);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mz_damp, dim3(cdiv(norbs, 256)), dim3(256), 0, s, norbs, idiag_d, W.dsc.as<double>(), p2_d,
                     p3_d, niter, three_point, pn_d, p1_d);
  MHIP_LAUNCH_CHECK();
  MHIP_CHECK(hipMemcpyAsync(p_d, pn_d, sizeof(double) * mpack, hipMemcpyDeviceToDevice, s));
  double sc[2];
  down(s, sc, W.dsc, sizeof(sc));
  MHIP_CHECK(hipStreamSynchronize(s));
  return sc[0];
}

void mz_diagg1_dev(hipStream_t s, const mhip_lmo_set* L, const double* fao_d, const double* p_d, double* eigs,
                   double* eigv, int idiagg, mhip_diagg1_state* st, int fmo_dim, double* fmo, int* ifmo, int* nij) {
  {
    MHIP_REQUIRE(st && fmo && ifmo && nij && eigs && eigv && fmo_dim > 0, "mhip_mozyme_diagg1: bad arguments");
    // saved state of diagg1.F90:61-71 (reset when st->reset is set)
    if (st->reset) {
      st->fref = st->oldens ? 0.0 : 10.0;
      st->safety = 1.0;
      st->oldlim = 0.0;
      st->reset = 0;
    }
    const bool full = (idiagg <= 5 || idiagg % 2 == 0);
    MHIP_REQUIRE(full || st->nfmo != nullptr, "mhip_mozyme_diagg1: nfmo needed to reuse the list");
    MzWork& W = mz_work();
    const LmoBufs& B = W.B;
    DevBuf &daocc = W.daocc, &dcounts = W.dcounts, &doffs = W.doffs,
           &dnfmo = W.dnfmo, &dold = W.dold, &dfmo = W.dfmo, &difmo = W.difmo, &dvmax = W.dvmax,
           &dvsum = W.dvsum, &deigv = W.deigv, &deigs = W.deigs;
    daocc.ensure(sizeof(double) * std::max(L->icocc_dim, 1));
    MHIP_CHECK(hipMemsetAsync(daocc.ptr, 0, sizeof(double) * std::max(L->icocc_dim, 1), s));
    up(s, deigv, eigv, sizeof(double) * std::max(L->nvir, 1));
    up(s, deigs, eigs, sizeof(double) * std::max(L->nocc, 1));
    D1Params P;
    const double cutlim = 1.0e-8;
    P.cutoff = std::max(cutlim, st->tiny * 10.0 * cutlim);
    P.flim = std::min(3.0, st->fref * 0.5);
    st->fref = 0.0;
    if (idiagg <= 5) P.cutoff = cutlim;
    P.oldlim = st->oldlim;
    P.full = full ? 1 : 0;
    P.fmo_dim = fmo_dim;
    P.fp32 = st->fp32_accumulate ? 1 : 0;
    const int nv = std::max(L->nvir, 1);
    dcounts.ensure(sizeof(int) * nv);
    doffs.ensure(sizeof(int) * nv);
    dnfmo.ensure(sizeof(int) * nv);
    dvmax.ensure(sizeof(double) * nv);
    dvsum.ensure(sizeof(double) * nv);
    dfmo.ensure(sizeof(double) * fmo_dim);
    difmo.ensure(sizeof(int) * 2 * fmo_dim);
    if (L->nocc > 0) {
      hipLaunchKernelGGL(k_mz_aocc, dim3(cdiv(L->nocc, 128)), dim3(128), 0, s, B.dev, daocc.as<double>());
      MHIP_LAUNCH_CHECK();
    }
    std::vector<int> offs(nv, 0), counts(nv, 0);
    CandDev C{};
    if (full && L->nocc > 0 && L->nocc <= 64 * MZ_CANDW) {
      const int na = L->numat;
      W.dfl_cnt.ensure(sizeof(int) * (na + 1));
      W.dfl_off.ensure(sizeof(int) * (na + 1));
      W.dfl_ent.ensure(sizeof(int) * 2 * L->nocc);
      hipLaunchKernelGGL(k_mz_zero_int, dim3(cdiv(na + 1, 256)), dim3(256), 0, s, na + 1, W.dfl_cnt.as<int>());
      MHIP_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_mz_first_count, dim3(cdiv(L->nocc, 256)), dim3(256), 0, s, B.dev, W.dfl_cnt.as<int>());
      MHIP_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_mz_scan_counts, dim3(1), dim3(1024), 0, s, na, W.dfl_cnt.as<int>(), W.dfl_off.as<int>());
      MHIP_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_mz_zero_int, dim3(cdiv(na + 1, 256)), dim3(256), 0, s, na + 1, W.dfl_cnt.as<int>());
      MHIP_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_mz_first_fill, dim3(cdiv(L->nocc, 256)), dim3(256), 0, s, B.dev, W.dfl_off.as<int>(),
                         W.dfl_cnt.as<int>(), W.dfl_ent.as<int>());
      MHIP_LAUNCH_CHECK();
      C.nbr_off = W.dnbr_off.as<int>();
      C.nbr = W.dnbr.as<int>();
      C.fl_off = W.dfl_off.as<int>();
      C.fl_ent = W.dfl_ent.as<int>();
      C.on = 1;
    }
    if (full) {
      if (L->nvir > 0) {
        hipLaunchKernelGGL(k_mz_diagg1_count, dim3(L->nvir), dim3(256), 0, s, B.dev, fao_d,
                           p_d, daocc.as<double>(), P, C, dcounts.as<int>(), deigv.as<double>());
        MHIP_LAUNCH_CHECK();
        down(s, counts.data(), dcounts, sizeof(int) * L->nvir);
        MHIP_CHECK(hipStreamSynchronize(s));
      }
      int acc = 0;
      for (int i = 0; i < L->nvir; ++i) {
        offs[i] = acc;
        acc += counts[i];
      }
    } else {
      // reuse: the old list's per-LMO counts (nfmo) and entries
      int acc = 0;
      for (int i = 0; i < L->nvir; ++i) {
        offs[i] = acc;
        acc += st->nfmo[i];
      }
      for (int k = 0; k < std::min(acc, fmo_dim); ++k)
        MHIP_REQUIRE(ifmo[2 * k] >= 1 && ifmo[2 * k] <= L->nvir && ifmo[2 * k + 1] >= 1 &&
                         ifmo[2 * k + 1] <= L->nocc,
                     "mhip_mozyme_diagg1: the list to reuse holds an entry out of range");
      up(s, dnfmo, st->nfmo, sizeof(int) * L->nvir);
      up(s, dold, ifmo, sizeof(int) * 2 * std::max(std::min(acc, fmo_dim), 1));
    }
    up(s, doffs, offs.data(), sizeof(int) * nv);
    if (L->nvir > 0) {
      hipLaunchKernelGGL(k_mz_diagg1_write, dim3(L->nvir), dim3(256), 0, s, B.dev, fao_d,
                         p_d, daocc.as<double>(), P, C, doffs.as<int>(), dnfmo.as<int>(),
                         full ? nullptr : dold.as<int>(), dfmo.as<double>(), difmo.as<int>(),
                         dvmax.as<double>(), dvsum.as<double>(), deigv.as<double>());
      MHIP_LAUNCH_CHECK();
    }
    std::vector<double> vmax(nv, 0.0), vsum(nv, 0.0);
    std::vector<int> nfmo_h(nv, 0);
    down(s, vmax.data(), dvmax, sizeof(double) * L->nvir);
    down(s, vsum.data(), dvsum, sizeof(double) * L->nvir);
    if (full) down(s, nfmo_h.data(), dnfmo, sizeof(int) * L->nvir);
    MHIP_CHECK(hipStreamSynchronize(s));
    double tiny = 0.0, sumt = 0.0;
    int total = 0;
    for (int i = 0; i < L->nvir; ++i) {
      tiny = std::max(tiny, vmax[i]);
      sumt += vsum[i];
      total += full ? nfmo_h[i] : st->nfmo[i];
    }
    total = std::min(total, fmo_dim);
    if (full && st->nfmo)
      for (int i = 0; i < L->nvir; ++i) st->nfmo[i] = nfmo_h[i];
    // storage full (diagg1.F90:458-470)
    if (total == fmo_dim) st->safety *= 2.0;
    else st->safety = std::max(st->safety * 0.5, 1.0);
    *nij = total;
    down(s, fmo, dfmo, sizeof(double) * total);
    down(s, ifmo, difmo, sizeof(int) * 2 * total);
    down(s, eigv, deigv, sizeof(double) * L->nvir);
    if (idiagg > 2 && idiagg % 4 != 0) {
      st->fref = (tiny * tiny) * (tiny * tiny);  // tiny**4 as the reference's powi
    } else {
      if (L->nocc > 0) {
        hipLaunchKernelGGL(k_mz_occ_energy, dim3(L->nocc), dim3(64), 0, s, B.dev, fao_d,
                           p_d, daocc.as<double>(), P.cutoff, deigs.as<double>());
        MHIP_LAUNCH_CHECK();
        down(s, eigs, deigs, sizeof(double) * L->nocc);
      }
      st->oldlim = tiny * st->safety * 1.0e-3;
      st->fref = (tiny * tiny) * (tiny * tiny);  // tiny**4 as the reference's powi
    }
    MHIP_CHECK(hipStreamSynchronize(s));
    st->tiny = tiny;
    st->sumt = sumt;
    st->ovmax = tiny;
  }
}

void mz_diagg2_dev(hipStream_t s, const mhip_lmo_set* L, const double* eigs, const double* eigv, const double* fmo,
                   const int* ifmo, int nij, int idiagg, mhip_diagg2_state* st) {
  {
    MHIP_REQUIRE(st && eigs && eigv && nij >= 0 && (nij == 0 || (fmo && ifmo)), "mhip_mozyme_diagg2: bad arguments");
    for (int k = 0; k < nij; ++k)
      MHIP_REQUIRE(ifmo[2 * k] >= 1 && ifmo[2 * k] <= L->nvir && ifmo[2 * k + 1] >= 1 &&
                       ifmo[2 * k + 1] <= L->nocc,
                   "mhip_mozyme_diagg2: ifmo entry out of range");
    if (st->reset) {
      st->nrejct[0] = st->nrejct[1] = 0;
      st->reset = 0;
    }
    D2Params P;
    // epseta (src/MOZYME/epseta.F90): eps = max(tiny(1d0), 1d-39); bigeps = 50 sqrt(eps)
    const double bigeps = 50.0 * std::sqrt(std::max(2.2250738585072014e-308, 1.0e-39));
    const bool retry = (st->nrejct[0] == st->nrejct[1] && st->nrejct[0] != 0 && st->nrejct[0] < 20);
    if (idiagg % 5 == 0 || idiagg <= 5) {
      st->tiny = -1.0;
      P.biglim = -1.0;
    } else {
      st->tiny = 0.01 * st->tiny;
      P.biglim = bigeps;
    }
    P.tiny = st->tiny;
    P.cnst = st->damp;
    P.shift = st->shift;
    P.thresh = st->thresh;
    P.retry = retry ? 1 : 0;
    P.nij = nij;
    P.lcap = std::min(g_mz_lcap, D2W_LCAP);
    P.pcap = std::min(g_mz_pcap, D2W_PCAP);
    // the rotations diagg2.F90 takes (same tests, same arithmetic as the
    // reference's loop), each at one level past the last earlier rotation
    // on its virtual or its occupied LMO
    std::vector<int> lev_v(std::max(L->nvir, 1), 0), lev_o(std::max(L->nocc, 1), 0), level;
    std::vector<int> taken;
    taken.reserve(nij);
    level.reserve(nij);
    int nlev = 0;
    // order-free: round = the first round in which neither LMO rotates yet
    // (bit r of used_v[i] / used_o[j], words per LMO grown as needed)
    int words = 4;
    std::vector<uint64_t> used_v, used_o;
    if (st->order_free) {
      used_v.assign((size_t)std::max(L->nvir, 1) * words, 0);
      used_o.assign((size_t)std::max(L->nocc, 1) * words, 0);
    }
    auto first_free = [&](int i, int j) {
      for (;;) {
        for (int w = 0; w < words; ++w) {
          const uint64_t f = ~(used_v[(size_t)i * words + w] | used_o[(size_t)j * words + w]);
          if (f) return w * 64 + __builtin_ctzll(f);
        }
        const int nw2 = 2 * words;
        auto grow = [&](std::vector<uint64_t>& u, int n) {
          std::vector<uint64_t> g((size_t)std::max(n, 1) * nw2, 0);
          for (int k = 0; k < n; ++k)
            for (int w = 0; w < words; ++w) g[(size_t)k * nw2 + w] = u[(size_t)k * words + w];
          u.swap(g);
        };
        grow(used_v, L->nvir);
        grow(used_o, L->nocc);
        words = nw2;
      }
    };
    for (int ij = 0; ij < nij; ++ij) {
      if (!(std::fabs(fmo[ij]) >= P.tiny)) continue;
      const int i = ifmo[2 * ij] - 1, j = ifmo[2 * ij + 1] - 1;
      const double c = fmo[ij] * P.cnst;
      const double d = eigs[j] - eigv[i] - P.shift;
      if (!(std::fabs(c / d) >= P.biglim)) continue;
      int l;
      if (st->order_free) {
        const int r = first_free(i, j);
        used_v[(size_t)i * words + r / 64] |= 1ull << (r % 64);
        used_o[(size_t)j * words + r / 64] |= 1ull << (r % 64);
        l = r + 1;
      } else {
        l = std::max(lev_v[i], lev_o[j]) + 1;
        lev_v[i] = lev_o[j] = l;
      }
      taken.push_back(ij);
      level.push_back(l);
      nlev = std::max(nlev, l);
    }
    // entries grouped by level (counting sort, list order inside a level)
    std::vector<int> start(nlev + 2, 0), order(taken.size());
    for (int l : level) ++start[l + 1];
    for (int l = 1; l <= nlev + 1; ++l) start[l] += start[l - 1];
    {
      std::vector<int> pos(start.begin(), start.end());
      for (size_t k = 0; k < taken.size(); ++k) order[pos[level[k]]++] = taken[k];
    }
    int widest = 0;
    for (int l = 1; l <= nlev; ++l) widest = std::max(widest, start[l + 1] - start[l]);
    // each LMO's coefficient span, as the kernel bounds it
    auto span_of = [&](int n, const int* no, int cdim) {
      int m = 1;
      for (int k = 0; k < n; ++k) {
        const int lo = no[k], hi = (k + 1 < n) ? no[k + 1] : cdim;
        m = std::max(m, std::min(lo + L->norbs, hi) - lo);
      }
      return m;
    };
    const int span_j = span_of(L->nocc, L->ncocc, L->cocc_dim);
    const int span_i = span_of(L->nvir, L->ncvir, L->cvir_dim);
    // The LDS kernels need 16-bit atom numbers and orbital offsets and their
    // maps in LDS.
    constexpr size_t kLds = 160 * 1024;
    const bool lds_ok = g_mz_one_wg && L->numat < (int)D2W_NONE && L->norbs < (int)D2W_NONE &&
                        d2w_lds_bytes(L->numat, 1) <= kLds;
    int nw = 0;
    while (lds_ok && nw < D2W_NW && d2w_lds_bytes(L->numat, nw + 1) <= kLds) ++nw;
    nw = std::min(nw, std::max(widest, 1));
    // a level wider than g_mz_wide_rounds rounds of the workgroup runs across
    // the chip, one workgroup per rotation; runs of narrower levels run in
    // one workgroup
    const int wide = lds_ok ? (int)std::min<long long>((long long)nw * g_mz_wide_rounds, 1LL << 30) : 0;
    int widest_wide = 0;
    if (lds_ok)
      for (int l = 1; l <= nlev; ++l)
        if (start[l + 1] - start[l] > wide) widest_wide = std::max(widest_wide, start[l + 1] - start[l]);
    const int nslots = lds_ok ? std::max(nw, std::min(widest_wide, 2048)) : std::max(1, std::min(widest, 2048));
    MzWork& W = mz_work();
    const LmoBufs& B = W.B;
    DevBuf &deigs = W.deigs, &deigv = W.deigv, &dfmo = W.dfmo, &difmo = W.difmo, &dorder = W.dorder,
           &dused = W.dused, &dlat = W.dlat, &dsj = W.dsj, &dsi = W.dsi, &dsumb = W.dsumb, &dnrej = W.dnrej;
    if (!lds_ok) {
      // the global-map kernel reads the list itself
      up(s, deigs, eigs, sizeof(double) * std::max(L->nocc, 1));
      up(s, deigv, eigv, sizeof(double) * std::max(L->nvir, 1));
      up(s, dfmo, fmo, sizeof(double) * std::max(nij, 1));
      up(s, difmo, ifmo, sizeof(int) * 2 * std::max(nij, 1));
      up(s, dorder, order.data(), sizeof(int) * order.size());
      dused.ensure(sizeof(int) * (size_t)nslots * L->numat);
      MHIP_CHECK(hipMemsetAsync(dused.ptr, 0xff, sizeof(int) * (size_t)nslots * L->numat, s));  // -1
      dlat.ensure((size_t)nslots * L->numat);
      MHIP_CHECK(hipMemsetAsync(dlat.ptr, 0, (size_t)nslots * L->numat, s));
    } else {
      // level l (0-based) holds the records [level[l], level[l + 1])
      up(s, W.dlevel, start.data() + 1, sizeof(int) * (size_t)(nlev + 1));
      // built straight into pinned memory (the previous call's copy out of it
      // ended at that call's final synchronization)
      D2Rot* recs = static_cast<D2Rot*>(W.hrec.ensure(sizeof(D2Rot) * std::max<size_t>(order.size(), 1)));
      for (size_t k = 0; k < order.size(); ++k) {
        const int ij = order[k], i = ifmo[2 * ij] - 1, j = ifmo[2 * ij + 1] - 1;
        D2Rot& r = recs[k];
        r.c = fmo[ij] * P.cnst;
        r.d = eigs[j] - eigv[i] - P.shift;
        r.ij = ij;
        r.i = i;
        r.j = j;
        r.nnj = L->nncf[j];
        r.nni = L->nnce[i];
        r.jlr = L->ncocc[j];
        r.ilr = L->ncvir[i];
        r.jur = std::min(r.jlr + L->norbs, (j + 1 < L->nocc) ? L->ncocc[j + 1] : L->cocc_dim);
        r.jncf = (j + 1 < L->nocc) ? L->nncf[j + 1] : L->icocc_dim;
        r.iur = std::min(r.ilr + L->norbs, (i + 1 < L->nvir) ? L->ncvir[i + 1] : L->cvir_dim);
        r.incv = (i + 1 < L->nvir) ? L->nnce[i + 1] : L->icvir_dim;
        r.pad = 0;
      }
      up(s, W.drec, recs, sizeof(D2Rot) * order.size());
    }
    dsj.ensure(sizeof(double) * (size_t)nslots * span_j);
    dsi.ensure(sizeof(double) * (size_t)nslots * span_i);
    dsumb.ensure(sizeof(double) * std::max(nij, 1));
    dnrej.ensure(sizeof(int) * std::max(nij, 1));
    if (g_mz_stats.on) {
      // algorithmic bytes: each rotation reads and writes the coefficients of
      // its two LMOs (their size on entry)
      auto coefs = [&](const int* nc, const int* nn, const int* idx, int k) {
        long long c = 0;
        for (int a = 0; a < nc[k]; ++a) c += L->nlast[idx[nn[k] + a] - 1] - L->nfirst[idx[nn[k] + a] - 1] + 1;
        return c;
      };
      for (int ij : taken) {
        const int i = ifmo[2 * ij] - 1, j = ifmo[2 * ij + 1] - 1;
        // a resident loop keeps each LMO's coefficient count from its last tidy
        g_mz_stats.bytes += 16.0 * (W.resident ? (double)W.nq_vir[i] + W.nq_occ[j]
                                               : (double)(coefs(L->nce, L->nnce, L->icvir, i) +
                                                          coefs(L->ncf, L->nncf, L->icocc, j)));
      }
      g_mz_stats.rotations += (long long)taken.size();
      g_mz_stats.levels += nlev;
      if (!g_mz_stats.e0) {
        MHIP_CHECK(hipEventCreate(&g_mz_stats.e0));
        MHIP_CHECK(hipEventCreate(&g_mz_stats.e1));
      }
      MHIP_CHECK(hipEventRecord(g_mz_stats.e0, s));
    }
    if (lds_ok && nlev > 0) {
      static bool attr_set = false;
      if (!attr_set) {
        MHIP_CHECK(hipFuncSetAttribute((const void*)k_mz_diagg2_wg, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)kLds));
        MHIP_CHECK(hipFuncSetAttribute((const void*)k_mz_diagg2_lvl, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)kLds));
        attr_set = true;
      }
      const size_t lds_wg = d2w_lds_bytes(L->numat, nw), lds_lvl = d2w_lds_bytes(L->numat, 1);
      int l = 1;
      while (l <= nlev) {
        const int width = start[l + 1] - start[l];
        if (width > wide) {
          for (int b = start[l]; b < start[l + 1]; b += nslots) {
            const int cnt = std::min(nslots, start[l + 1] - b);
            hipLaunchKernelGGL(k_mz_diagg2_lvl, dim3(cnt), dim3(64), lds_lvl, s, B.dev, W.drec.as<D2Rot>(), P, b,
                               dsj.as<double>(), span_j, dsi.as<double>(), span_i, dsumb.as<double>(),
                               dnrej.as<int>());
            MHIP_LAUNCH_CHECK();
            if (g_mz_stats.on) {
              ++g_mz_stats.launches;
              ++g_mz_stats.rounds;
            }
          }
          ++l;
          continue;
        }
        int l1 = l;  // the run of narrow levels l .. l1 - 1
        while (l1 <= nlev && start[l1 + 1] - start[l1] <= wide) ++l1;
        hipLaunchKernelGGL(k_mz_diagg2_wg, dim3(1), dim3(64 * nw), lds_wg, s, B.dev, W.drec.as<D2Rot>(), P,
                           W.dlevel.as<int>(), l - 1, l1 - 1, dsj.as<double>(), span_j, dsi.as<double>(), span_i,
                           dsumb.as<double>(), dnrej.as<int>());
        MHIP_LAUNCH_CHECK();
        if (g_mz_stats.on) {
          ++g_mz_stats.launches;
          for (int q = l; q < l1; ++q) g_mz_stats.rounds += (start[q + 1] - start[q] + nw - 1) / nw;
        }
        l = l1;
      }
    }
    for (int l = 1; l <= nlev && !lds_ok; ++l)
      for (int b = start[l]; b < start[l + 1]; b += nslots) {
        const int cnt = std::min(nslots, start[l + 1] - b);
        hipLaunchKernelGGL(k_mz_diagg2, dim3(cnt), dim3(D2_NT), 0, s, B.dev, deigs.as<double>(),
                           deigv.as<double>(), dfmo.as<double>(), difmo.as<int>(), P, dorder.as<int>(), b,
                           dused.as<int>(), (char*)dlat.ptr, dsj.as<double>(), span_j, dsi.as<double>(), span_i,
                           dsumb.as<double>(), dnrej.as<int>());
        MHIP_LAUNCH_CHECK();
        if (g_mz_stats.on) {
          ++g_mz_stats.launches;
          ++g_mz_stats.rounds;
        }
      }
    if (g_mz_stats.on) MHIP_CHECK(hipEventRecord(g_mz_stats.e1, s));
    const double* sb = static_cast<const double*>(W.hsumb.ensure(sizeof(double) * std::max(nij, 1)));
    const int* nr = static_cast<const int*>(W.hnrej.ensure(sizeof(int) * std::max(nij, 1)));
    if (!taken.empty()) {
      down(s, W.hsumb.ptr, dsumb, sizeof(double) * nij);
      down(s, W.hnrej.ptr, dnrej, sizeof(int) * nij);
    }
    MHIP_CHECK(hipStreamSynchronize(s));
    double sumb = 0.0;
    int nrej = 0;
    for (int ij : taken) {  // list order
      sumb += sb[ij];
      nrej += nr[ij];
    }
    st->sumb = sumb;
    st->nrejct[1] = st->nrejct[0];
    st->nrejct[0] = nrej;
    if (g_mz_stats.on) {
      float t = 0.f;
      MHIP_CHECK(hipEventElapsedTime(&t, g_mz_stats.e0, g_mz_stats.e1));
      g_mz_stats.ms += t;
    }
  }
}


hipStream_t mz_stream_of(mhip_ctx* ctx) { return mz_stream(ctx); }

// The LMO set uploaded once (checked as every entry point checks it) and
// owned by the device until mz_resident_end copies it back.
void mz_resident_begin(hipStream_t s, const mhip_lmo_set* L) {
  MzWork& W = mz_work();
  upload_lmos(s, L, W);
  W.resident = true;
  auto counts = [&](int n, const int* nc, const int* nn, const int* idx, std::vector<int>& out) {
    out.assign(std::max(n, 1), 0);
    for (int k = 0; k < n; ++k)
      for (int a = 0; a < nc[k]; ++a) out[k] += L->nlast[idx[nn[k] + a] - 1] - L->nfirst[idx[nn[k] + a] - 1] + 1;
  };
  counts(L->nocc, L->ncf, L->nncf, L->icocc, W.nq_occ);
  counts(L->nvir, L->nce, L->nnce, L->icvir, W.nq_vir);
}

// Everything tidy and diagg2 change, back into the caller's arrays.
void mz_resident_end(hipStream_t s, mhip_lmo_set* L) {
  MzWork& W = mz_work();
  const LmoBufs& B = W.B;
  W.resident = false;
  down(s, L->ncf, B.ncf, sizeof(int) * L->nocc);
  down(s, const_cast<int*>(L->nncf), B.nncf, sizeof(int) * L->nocc);
  down(s, const_cast<int*>(L->ncocc), B.ncocc, sizeof(int) * L->nocc);
  down(s, L->icocc, B.icocc, sizeof(int) * L->icocc_dim);
  down(s, L->cocc, B.cocc, sizeof(double) * L->cocc_dim);
  down(s, L->nce, B.nce, sizeof(int) * L->nvir);
  down(s, const_cast<int*>(L->nnce), B.nnce, sizeof(int) * L->nvir);
  down(s, const_cast<int*>(L->ncvir), B.ncvir, sizeof(int) * L->nvir);
  down(s, L->icvir, B.icvir, sizeof(int) * L->icvir_dim);
  down(s, L->cvir, B.cvir, sizeof(double) * L->cvir_dim);
  MHIP_CHECK(hipStreamSynchronize(s));
}

// tidy (src/MOZYME/tidy.F90:16-244) at a single point on the resident
// occupied (occ) or virtual set: the per-LMO passes on the device, the
// free-space spreading (an order-dependent recurrence over the LMO counts)
// on the host, which also keeps the new offsets in nn_host / no_host.
// Returns false when the storage is too small for the next annihilation
// step (the reference then re-allocates 60% larger arrays).
bool mz_tidy_dev(hipStream_t s, int numat, int norbs, bool occ, int nmo, int* nn_host, int* no_host,
                 double thresh) {
  MzWork& W = mz_work();
  LmoBufs& B = W.B;
  if (nmo == 0) return true;
  DevBuf& nc = occ ? B.ncf : B.nce;
  DevBuf& nnc = occ ? B.nncf : B.nnce;
  DevBuf& ncmo = occ ? B.ncocc : B.ncvir;
  DevBuf& ic = occ ? B.icocc : B.icvir;
  DevBuf& c = occ ? B.cocc : B.cvir;
  DevBuf& ic2 = occ ? W.icocc2 : W.icvir2;
  DevBuf& c2 = occ ? W.cocc2 : W.cvir2;
  const int n01 = occ ? B.dev.icocc_dim : B.dev.icvir_dim;
  const int n02 = occ ? B.dev.cocc_dim : B.dev.cvir_dim;
  W.dkeep.ensure((size_t)n01);
  W.dkc.ensure(sizeof(int) * nmo);
  W.dkq.ensure(sizeof(int) * nmo);
  hipLaunchKernelGGL(k_mz_tidy_count, dim3(cdiv(nmo, 128)), dim3(128), 0, s, nmo, nc.as<int>(), nnc.as<int>(),
                     ncmo.as<int>(), ic.as<int>(), c.as<double>(), B.iorbs.as<int>(), thresh,
                     (unsigned char*)W.dkeep.ptr, W.dkc.as<int>(), W.dkq.as<int>());
  MHIP_LAUNCH_CHECK();
  std::vector<int> kc(nmo), kq(nmo);
  down(s, kc.data(), W.dkc, sizeof(int) * nmo);
  down(s, kq.data(), W.dkq, sizeof(int) * nmo);
  MHIP_CHECK(hipStreamSynchronize(s));
  long long ln = 0, mn = 0;
  for (int i = 0; i < nmo; ++i) {
    ln += kc[i];
    mn += kq[i];
  }
  int ispace = (int)((n01 - ln) / nmo), jspace = (int)((n02 - mn) / nmo);
  const int per = n01 / nmo;
  if (per + ispace < numat && ispace < std::max(per / 5, 20)) return false;
  if (ispace > numat && jspace > nmo) ispace = jspace = 0;
  std::vector<int> nn_new(nmo), no_new(nmo);
  int jtop = n01, mtop = n02;
  int jsav = (int)(n01 - ln - (long long)ispace * nmo), msav = (int)(n02 - mn - (long long)jspace * nmo);
  long long lc = ln - kc[nmo - 1], mc = mn - kq[nmo - 1];  // compacted starts of LMO i
  for (int i = nmo - 1; i >= 1; --i) {
    int j = std::min(kc[i] + ispace, numat);
    if (j < kc[i] + ispace) {
      jsav += kc[i] + ispace - j;
      jtop -= j;
    } else {
      const int jdash = std::min(numat, std::max(jsav, 0) + j);
      jtop -= jdash;
      jsav = jsav - jdash + j;
    }
    // the moved LMO lands at or above its compacted place (never over an
    // LMO below it); the reference's bookkeeping keeps it so
    if (jtop < lc) return false;
    nn_new[i] = jtop;
    const int n = kq[i];
    j = std::min(n + jspace, norbs);
    if (j < n + jspace) {
      msav += n + jspace - j;
      mtop -= j;
    } else {
      const int mdash = std::min(norbs, std::max(msav, 0) + j);
      mtop -= mdash;
      msav = msav - mdash + j;
    }
    if (mtop < mc) return false;
    no_new[i] = mtop;
    lc -= kc[i - 1];
    mc -= kq[i - 1];
  }
  nn_new[0] = 0;
  no_new[0] = 0;
  up(s, W.dnn_new, nn_new.data(), sizeof(int) * nmo);
  up(s, W.dnq_new, no_new.data(), sizeof(int) * nmo);
  ic2.ensure(sizeof(int) * (size_t)n01);
  c2.ensure(sizeof(double) * (size_t)n02);
  hipLaunchKernelGGL(k_mz_tidy_move, dim3(cdiv(nmo, 128)), dim3(128), 0, s, nmo, nc.as<int>(), nnc.as<int>(),
                     ncmo.as<int>(), ic.as<int>(), c.as<double>(), B.iorbs.as<int>(),
                     (const unsigned char*)W.dkeep.ptr, W.dnn_new.as<int>(), W.dnq_new.as<int>(), ic2.as<int>(),
                     c2.as<double>());
  MHIP_LAUNCH_CHECK();
  MHIP_CHECK(hipMemcpyAsync(nnc.ptr, W.dnn_new.ptr, sizeof(int) * nmo, hipMemcpyDeviceToDevice, s));
  MHIP_CHECK(hipMemcpyAsync(ncmo.ptr, W.dnq_new.ptr, sizeof(int) * nmo, hipMemcpyDeviceToDevice, s));
  std::swap(ic, ic2);
  std::swap(c, c2);
  LmoDev& d = B.dev;
  if (occ) {
    d.icocc = B.icocc.as<int>();
    d.cocc = B.cocc.as<double>();
    W.nq_occ = kq;
  } else {
    d.icvir = B.icvir.as<int>();
    d.cvir = B.cvir.as<double>();
    W.nq_vir = kq;
  }
  std::copy(nn_new.begin(), nn_new.end(), nn_host);
  std::copy(no_new.begin(), no_new.end(), no_host);
  return true;
}

// check (src/MOZYME/check.F90:16-113) on the resident set: renormalizes every
// LMO and returns the summed normalization error (the reference stops the
// job above 0.1).
double mz_check_dev(hipStream_t s, bool occ, int nmo) {
  MzWork& W = mz_work();
  LmoBufs& B = W.B;
  if (nmo == 0) return 0.0;
  W.derr.ensure(sizeof(double) * nmo);
  hipLaunchKernelGGL(k_mz_check, dim3(cdiv(nmo, 128)), dim3(128), 0, s, nmo,
                     (occ ? B.ncf : B.nce).as<int>(), (occ ? B.nncf : B.nnce).as<int>(),
                     (occ ? B.ncocc : B.ncvir).as<int>(), (occ ? B.icocc : B.icvir).as<int>(),
                     (occ ? B.cocc : B.cvir).as<double>(), B.iorbs.as<int>(), W.derr.as<double>());
  MHIP_LAUNCH_CHECK();
  std::vector<double> err(nmo);
  down(s, err.data(), W.derr, sizeof(double) * nmo);
  MHIP_CHECK(hipStreamSynchronize(s));
  double e = 0.0;
  for (double x : err) e += x;
  return e;
}

// The stored off-diagonal blocks of nijbo (offset, length), for eimp.
void mz_eimp_blocks(hipStream_t s, int numat, const int* nijbo, const int* iorbs, int* nblk) {
  MzWork& W = mz_work();
  std::vector<int> off, len;
  for (int i = 0; i < numat; ++i)
    for (int j = 0; j < i; ++j) {
      const int k = nijbo[i + (size_t)j * numat];
      if (k >= 0) {
        off.push_back(k);
        len.push_back(iorbs[i] * iorbs[j]);
      }
    }
  *nblk = (int)off.size();
  up(s, W.deoff, off.data(), sizeof(int) * off.size());
  up(s, W.delen, len.data(), sizeof(int) * len.size());
}

void mz_eimp_dev(hipStream_t s, int nblk, const double* f_d, double* p_d) {
  MzWork& W = mz_work();
  if (nblk == 0) return;
  hipLaunchKernelGGL(k_mz_eimp, dim3(cdiv(nblk, 256)), dim3(256), 0, s, nblk, W.deoff.as<int>(),
                     W.delen.as<int>(), f_d, p_d);
  MHIP_LAUNCH_CHECK();
}

}  // namespace mhip

extern "C" {

int mhip_mozyme_density(mhip_ctx* ctx, const mhip_lmo_set* L, int mode, const double* partp, double* p) {
  return mz_guard([&] {
    hipStream_t s = mz_stream(ctx);
    MzWork& W = mz_work();
    upload_lmos(s, L, W);
    DevBuf &dp = W.dp, &dpart = W.dpart;
    dp.ensure(sizeof(double) * L->mpack);
    if (mode != 0) {
      MHIP_REQUIRE(partp != nullptr, "mhip_mozyme_density: partp required for mode != 0");
      up(s, dpart, partp, sizeof(double) * L->mpack);
    }
    mhip::mz_density_dev(s, L->mpack, L->nocc, mode, mode != 0 ? dpart.as<double>() : nullptr, dp.as<double>());
    down(s, p, dp, sizeof(double) * L->mpack);
    MHIP_CHECK(hipStreamSynchronize(s));
  });
}

int mhip_mozyme_diagg1(mhip_ctx* ctx, const mhip_lmo_set* L, const double* fao, const double* p,
                       double* eigs, double* eigv, int idiagg, mhip_diagg1_state* st, int fmo_dim,
                       double* fmo, int* ifmo, int* nij) {
  return mz_guard([&] {
    hipStream_t s = mz_stream(ctx);
    MzWork& W = mz_work();
    upload_lmos(s, L, W);
    up(s, W.dfao, fao, sizeof(double) * L->mpack);
    up(s, W.dp, p, sizeof(double) * L->mpack);
    mhip::mz_diagg1_dev(s, L, W.dfao.as<double>(), W.dp.as<double>(), eigs, eigv, idiagg, st, fmo_dim, fmo, ifmo,
                        nij);
  });
}

int mhip_mozyme_diagg2(mhip_ctx* ctx, mhip_lmo_set* L, const double* eigs, const double* eigv,
                       const double* fmo, const int* ifmo, int nij, int idiagg, mhip_diagg2_state* st) {
  return mz_guard([&] {
    hipStream_t s = mz_stream(ctx);
    MzWork& W = mz_work();
    upload_lmos(s, L, W);
    mhip::mz_diagg2_dev(s, L, eigs, eigv, fmo, ifmo, nij, idiagg, st);
    const LmoBufs& B = W.B;
    down(s, L->ncf, B.ncf, sizeof(int) * L->nocc);
    down(s, L->icocc, B.icocc, sizeof(int) * L->icocc_dim);
    down(s, L->cocc, B.cocc, sizeof(double) * L->cocc_dim);
    down(s, L->nce, B.nce, sizeof(int) * L->nvir);
    down(s, L->icvir, B.icvir, sizeof(int) * L->icvir_dim);
    down(s, L->cvir, B.cvir, sizeof(double) * L->cvir_dim);
    MHIP_CHECK(hipStreamSynchronize(s));
  });
}

int mhip_mozyme_stats(int enable, double* ms, long long* rotations, long long* levels, long long* launches,
                      double* bytes, long long* rounds) {
  if (ms) *ms = g_mz_stats.ms;
  if (rotations) *rotations = g_mz_stats.rotations;
  if (levels) *levels = g_mz_stats.levels;
  if (launches) *launches = g_mz_stats.launches;
  if (bytes) *bytes = g_mz_stats.bytes;
  if (rounds) *rounds = g_mz_stats.rounds;
  if (enable >= 0) {
    g_mz_stats.on = enable != 0;
    g_mz_stats.ms = g_mz_stats.bytes = 0.0;
    g_mz_stats.rotations = g_mz_stats.levels = g_mz_stats.launches = g_mz_stats.rounds = 0;
  }
  return 0;
}

int mhip_mozyme_tidy(mhip_ctx* ctx, mhip_lmo_set* L, int occupied, double thresh) {
  return mz_guard([&] {
    MHIP_REQUIRE(L, "mhip_mozyme_tidy: null LMO set");
    hipStream_t s = mz_stream(ctx);
    mhip::mz_resident_begin(s, L);
    const bool occ = occupied != 0;
    const int nmo = occ ? L->nocc : L->nvir;
    std::vector<int> nn(std::max(nmo, 1)), no(std::max(nmo, 1));
    const bool ok = mhip::mz_tidy_dev(s, L->numat, L->norbs, occ, nmo, nn.data(), no.data(), thresh);
    mhip::mz_resident_end(s, L);
    MHIP_REQUIRE(ok, "mhip_mozyme_tidy: LMO storage too small for the next annihilation step");
  });
}

int mhip_mozyme_check(mhip_ctx* ctx, mhip_lmo_set* L, int occupied, double* error) {
  return mz_guard([&] {
    MHIP_REQUIRE(L && error, "mhip_mozyme_check: bad arguments");
    hipStream_t s = mz_stream(ctx);
    mhip::mz_resident_begin(s, L);
    *error = mhip::mz_check_dev(s, occupied != 0, occupied ? L->nocc : L->nvir);
    mhip::mz_resident_end(s, L);
  });
}

int mhip_mozyme_eimp(mhip_ctx* ctx, int numat, const int* nfirst, const int* nlast, const int* nijbo,
                     long long mpack, const double* f, double* p) {
  return mz_guard([&] {
    MHIP_REQUIRE(numat > 0 && nfirst && nlast && nijbo && f && p, "mhip_mozyme_eimp: bad arguments");
    hipStream_t s = mz_stream(ctx);
    std::vector<int> iorbs(numat);
    for (int a = 0; a < numat; ++a) iorbs[a] = nlast[a] - nfirst[a] + 1;
    MzWork& W = mz_work();
    upload_nijbo(s, W, numat, nijbo, mpack, iorbs);
    int nblk = 0;
    mhip::mz_eimp_blocks(s, numat, nijbo, iorbs.data(), &nblk);
    up(s, W.df, f, sizeof(double) * mpack);
    up(s, W.dp, p, sizeof(double) * mpack);
    mhip::mz_eimp_dev(s, nblk, W.df.as<double>(), W.dp.as<double>());
    down(s, p, W.dp, sizeof(double) * mpack);
    MHIP_CHECK(hipStreamSynchronize(s));
  });
}

int mhip_mozyme_diagg2_one_workgroup(int enable) {
  const int was = g_mz_one_wg ? 1 : 0;
  if (enable >= 0) g_mz_one_wg = enable != 0;
  return was;
}

int mhip_mozyme_diagg2_wide_rounds(int rounds) {
  const int was = g_mz_wide_rounds;
  if (rounds >= 0) g_mz_wide_rounds = rounds;
  return was;
}

int mhip_mozyme_diagg2_list_cap(int cap) {
  const int was = g_mz_lcap;
  if (cap >= 0) g_mz_lcap = std::min(cap, D2W_LCAP);
  return was;
}

int mhip_mozyme_diagg2_coef_cap(int cap) {
  const int was = g_mz_pcap;
  if (cap >= 0) g_mz_pcap = std::min(cap, D2W_PCAP);
  return was;
}

int mhip_mozyme_helecz(mhip_ctx* ctx, int numat, const int* nfirst, const int* nlast, const int* nijbo,
                       long long mpack, const double* p, const double* h, const double* f, double* e) {
  return mz_guard([&] {
    MHIP_REQUIRE(numat > 0 && nfirst && nlast && nijbo && p && h && f && e, "mhip_mozyme_helecz: bad arguments");
    hipStream_t s = mz_stream(ctx);
    std::vector<int> iorbs(numat);
    for (int a = 0; a < numat; ++a) iorbs[a] = nlast[a] - nfirst[a] + 1;
    MzWork& W = mz_work();
    upload_nijbo(s, W, numat, nijbo, mpack, iorbs);
    up(s, W.dp, p, sizeof(double) * mpack);
    up(s, W.dh, h, sizeof(double) * mpack);
    up(s, W.df, f, sizeof(double) * mpack);
    *e = mhip::mz_helecz_dev(s, numat, W.dp.as<double>(), W.dh.as<double>(), W.df.as<double>());
  });
}

int mhip_mozyme_cnvgz(mhip_ctx* ctx, int norbs, long long mpack, double* pnew, double* p, double* p1,
                      double* p2, double* p3, int niter, const int* idiag, int three_point, double* pmax) {
  return mz_guard([&] {
    MHIP_REQUIRE(norbs > 0 && mpack > 0 && pnew && p && p1 && p2 && p3 && idiag && pmax,
                 "mhip_mozyme_cnvgz: bad arguments");
    for (int i = 0; i < norbs; ++i)
      MHIP_REQUIRE(idiag[i] >= 1 && idiag[i] <= mpack, "mhip_mozyme_cnvgz: idiag out of range");
    hipStream_t s = mz_stream(ctx);
    MzWork& W = mz_work();
    DevBuf &dpn = W.dpn, &dp = W.dp, &d1 = W.d1, &d2 = W.d2, &d3 = W.d3, &di = W.di;
    up(s, dpn, pnew, sizeof(double) * mpack);
    up(s, dp, p, sizeof(double) * mpack);
    up(s, d1, p1, sizeof(double) * norbs);
    up(s, d2, p2, sizeof(double) * norbs);
    up(s, d3, p3, sizeof(double) * norbs);
    up(s, di, idiag, sizeof(int) * norbs);
    *pmax = mhip::mz_cnvgz_dev(s, norbs, mpack, dpn.as<double>(), dp.as<double>(), d1.as<double>(),
                               d2.as<double>(), d3.as<double>(), niter, di.as<int>(), three_point);
    down(s, pnew, dpn, sizeof(double) * mpack);
    down(s, p, dpn, sizeof(double) * mpack);  // p = pnew
    down(s, p1, d1, sizeof(double) * norbs);
    down(s, p2, d2, sizeof(double) * norbs);
    down(s, p3, d3, sizeof(double) * norbs);
    MHIP_CHECK(hipStreamSynchronize(s));
  });
}

}  // extern "C"
