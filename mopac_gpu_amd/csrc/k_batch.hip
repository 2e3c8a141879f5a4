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
// molecule: warm-started cyclic Jacobi in parallel order, matrix in global
// memory/L2) -> cnvg (batched) -> control (PL, frontier levels, newdg) ->
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
};

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
  if (tid == 0) st[m].have_c = 1;
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
  std::vector<int> cls_count;
  DevBuf cls_list;
  int tile[4] = {64, 128, 256, 256};   // workgroup size per class (n <= 32, 64, 128, 256)
  double tol = 1.0e-10;
  std::vector<BatchState> init_state;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_iters = 0;
};

namespace {

int size_class(int n) { return n <= 32 ? 0 : (n <= 64 ? 1 : (n <= 128 ? 2 : 3)); }

template <int NT>
void launch_beig(hipStream_t s, int nblk, const BatchMol* mols, const int* doeig, BatchState* st, const int* list,
                 double tol) {
  hipLaunchKernelGGL(k_beig<NT>, dim3(nblk), dim3(NT), 0, s, mols, doeig, st, list, tol);
  MHIP_LAUNCH_CHECK();
}

void launch_eig_class(mhip_batch* b, int cls, int offset, int count) {
  if (count <= 0) return;
  const int* list = b->cls_list.as<int>() + offset;
  auto* mols = b->mols.as<BatchMol>();
  auto* st = b->st.as<BatchState>();
  const int* de = b->doeig.as<int>();
  switch (b->tile[cls]) {
    case 64: launch_beig<64>(b->s, count, mols, de, st, list, b->tol); break;
    case 128: launch_beig<128>(b->s, count, mols, de, st, list, b->tol); break;
    case 512: launch_beig<512>(b->s, count, mols, de, st, list, b->tol); break;
    default: launch_beig<256>(b->s, count, mols, de, st, list, b->tol); break;
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
    launch_eig_class(b, c, off, b->cls_count[c]);
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
    std::vector<long long> o_w(nmol), o_mp(nmol), o_nn(nmol), o_n(nmol), o_at(nmol), o_cc(nmol);
    long long tw = 0, tmp = 0, tnn = 0, tn = 0, tat = 0, tcc = 0;
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
    for (DevBuf* x : {&b->C, &b->A, &b->T}) x->ensure(sizeof(double) * tnn);
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
      q.A = b->A.as<double>() + o_nn[m];
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
    for (int c = 0; c < 4; ++c)
      for (int m = 0; m < nmol; ++m)
        if (size_class(b->n_h[m]) == c) {
          order.push_back(m);
          b->cls_count[c]++;
        }
    upv(b->cls_list, order.data(), sizeof(int) * nmol);
    MHIP_CHECK(hipStreamSynchronize(s));
    *out = b;
  });
}

int mhip_batch_set_tiles(mhip_batch* b, const int* tiles) {
  return bguard([&] {
    MHIP_REQUIRE(b && tiles, "mhip_batch_set_tiles: bad arguments");
    for (int c = 0; c < 4; ++c) {
      MHIP_REQUIRE(tiles[c] == 64 || tiles[c] == 128 || tiles[c] == 256 || tiles[c] == 512,
                   "mhip_batch_set_tiles: workgroup sizes are 64, 128, 256 or 512");
      b->tile[c] = tiles[c];
    }
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
    }
  });
}

// Run the batched SCF from the starting densities: capture `iters`
// iterations once into a graph, replay it until every molecule is done.
int mhip_batch_run(mhip_batch* b, int iters_per_graph, int max_launches, double* seconds, int* n_converged) {
  return bguard([&] {
    MHIP_REQUIRE(b && iters_per_graph > 0 && max_launches > 0, "mhip_batch_run: bad arguments");
    hipStream_t s = b->s;
    // reset the per-molecule state and the starting densities (outside the graph)
    MHIP_CHECK(hipMemcpyAsync(b->st.ptr, b->init_state.data(), sizeof(BatchState) * b->nmol,
                              hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(b->p.ptr, b->p0.data(), sizeof(double) * b->p0.size(), hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(b->pa.ptr, b->pa0.data(), sizeof(double) * b->pa0.size(), hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(b->pold.ptr, b->pa0.data(), sizeof(double) * b->pa0.size(), hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(b->p1.ptr, b->p1h.data(), sizeof(double) * b->p1h.size(), hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipStreamSynchronize(s));
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
