/*
 * CPU FP64 restatement of MOPAC's RHF SCF inner loop — TEST INFRASTRUCTURE.
 *
 * This file is the oracle against which the HIP path is checked.  Only the
 * tests, the smoke check and the bench's cpu_baseline leg load it; the product
 * library never links, calls or falls back to it.
 *
 * Each routine follows the reference routine cited beside it (paths relative
 * to the reference root).  Packed storage is MOPAC's lower triangle by rows:
 * element (i,j), i>=j, 1-based, at i(i-1)/2 + j.
 *
 * The dense eigensolver is LAPACK dsyevd, as in the reference
 * (src/matrix/eigenvectors_LAPACK.F90:96-120); it is taken from MKL
 * (libmkl_rt, LP64) when that library can be opened, and otherwise from a
 * cyclic Jacobi solver written here (adequate for the small test systems).
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PK(i, j) ((i) >= (j) ? ((long long)(i) * ((i) - 1) / 2 + (j)) : ((long long)(j) * ((j) - 1) / 2 + (i)))

/* ------------------------------------------------------------------------ */
/* LAPACK binding                                                            */
/* ------------------------------------------------------------------------ */
typedef void (*dsyevd_fn)(const char*, const char*, const int*, double*, const int*, double*,
                          double*, const int*, int*, const int*, int*);
typedef void (*dsyrk_fn)(const char*, const char*, const int*, const int*, const double*,
                         const double*, const int*, const double*, double*, const int*);
typedef void (*dsymm_fn)(const char*, const char*, const int*, const int*, const double*,
                         const double*, const int*, const double*, const int*, const double*,
                         double*, const int*);
typedef int (*mkl_set_threads_fn)(int);
static dsyevd_fn p_dsyevd = NULL;
static dsyrk_fn p_dsyrk = NULL;
static dsymm_fn p_dsymm = NULL;
static mkl_set_threads_fn p_set_threads = NULL;
static int g_lapack_tried = 0;

int o_lapack_available(void) {
  if (!g_lapack_tried) {
    g_lapack_tried = 1;
    const char* cands[] = {"/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1",
                           "libmkl_rt.so", NULL};
    for (int i = 0; cands[i]; ++i) {
      void* hdl = dlopen(cands[i], RTLD_NOW | RTLD_GLOBAL);
      if (!hdl) continue;
      p_dsyevd = (dsyevd_fn)dlsym(hdl, "dsyevd_");
      p_dsyrk = (dsyrk_fn)dlsym(hdl, "dsyrk_");
      p_dsymm = (dsymm_fn)dlsym(hdl, "dsymm_");
      p_set_threads = (mkl_set_threads_fn)dlsym(hdl, "MKL_Set_Num_Threads");
      if (!p_set_threads) p_set_threads = (mkl_set_threads_fn)dlsym(hdl, "mkl_set_num_threads");
      if (p_dsyevd) break;
    }
  }
  return p_dsyevd != NULL;
}

void o_set_threads(int n) {
  if (o_lapack_available() && p_set_threads) p_set_threads(n);
}

/* Cyclic Jacobi for a dense symmetric matrix (column-major, n x n).  On exit
 * a holds the eigenvectors (columns) sorted by ascending eigenvalue. */
static void jacobi_eigh(int n, double* a, double* w) {
  double* v = (double*)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double x = a[(size_t)j * n + i];
        tot += x * x;
        if (i != j) off += x * x;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[(size_t)q * n + p];
        if (fabs(apq) < 1e-300) continue;
        double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        double theta = 0.5 * (aqq - app) / apq;
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = a[(size_t)p * n + k], akq = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * akp - s * akq;
          a[(size_t)q * n + k] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = a[(size_t)k * n + p], aqk = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * apk - s * aqk;
          a[(size_t)k * n + q] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = v[(size_t)p * n + k], vkq = v[(size_t)q * n + k];
          v[(size_t)p * n + k] = c * vkp - s * vkq;
          v[(size_t)q * n + k] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
  /* selection sort by eigenvalue */
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (w[j] < w[m]) m = j;
    if (m != i) {
      double t = w[i]; w[i] = w[m]; w[m] = t;
      for (int k = 0; k < n; ++k) {
        double x = v[(size_t)i * n + k];
        v[(size_t)i * n + k] = v[(size_t)m * n + k];
        v[(size_t)m * n + k] = x;
      }
    }
  }
  memcpy(a, v, sizeof(double) * (size_t)n * n);
  free(v);
}

/* Dense symmetric eigensolve of an n x n column-major upper-stored matrix. */
int o_dsyevd(int n, double* a, double* w) {
  if (o_lapack_available()) {
    int lwork = -1, liwork = -1, info = 0;
    double wq;
    int iwq;
    p_dsyevd("V", "U", &n, a, &n, w, &wq, &lwork, &iwq, &liwork, &info);
    lwork = (int)wq;
    liwork = iwq;
    double* work = (double*)malloc(sizeof(double) * (size_t)lwork);
    int* iwork = (int*)malloc(sizeof(int) * (size_t)liwork);
    p_dsyevd("V", "U", &n, a, &n, w, work, &lwork, iwork, &liwork, &info);
    free(work);
    free(iwork);
    return info;
  }
  /* symmetrize from the upper triangle, then Jacobi */
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) a[(size_t)j * n + i] = a[(size_t)i * n + j];
  jacobi_eigh(n, a, w);
  return 0;
}

/* eigenvectors_LAPACK (src/matrix/eigenvectors_LAPACK.F90:16-127): adds
 * i*1e-10 to the diagonal of the packed matrix IN PLACE, unpacks the upper
 * triangle (dtpttr 'U'), and calls dsyevd('V','U'). */
int o_eigenvectors(int n, double* c, double* fpacked, double* eig) {
  for (int i = 1; i <= n; ++i) fpacked[PK(i, i) - 1] += i * 1.0e-10;
  memset(c, 0, sizeof(double) * (size_t)n * n);
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= j; ++i) c[(size_t)(j - 1) * n + (i - 1)] = fpacked[PK(j, i) - 1];
  return o_dsyevd(n, c, eig);
}

/* phase_lock (src/forces/force.F90:651-680): largest-|c| coefficient positive. */
void o_phase_lock(int n, double* c) {
  for (int i = 0; i < n; ++i) {
    double asum = 0.0, ssum = 0.0;
    for (int j = 0; j < n; ++j) {
      double x = c[(size_t)i * n + j];
      if (fabs(x) > asum) { asum = fabs(x); ssum = x; }
    }
    if (ssum < 0.0)
      for (int j = 0; j < n; ++j) c[(size_t)i * n + j] = -c[(size_t)i * n + j];
  }
}

/* ------------------------------------------------------------------------ */
/* Density (src/matrix/density_for_GPU.F90:100-108, case 5, fract < 1e-2):    */
/*   P = occ * C(:,1:ndubl) C(:,1:ndubl)^T, packed lower triangle.            */
/* ------------------------------------------------------------------------ */
void o_density(int n, int ndubl, double occ, const double* c, double* p) {
  if (o_lapack_available() && p_dsyrk && n > 64) {
    double* x = (double*)calloc((size_t)n * n, sizeof(double));
    double zero = 0.0;
    p_dsyrk("U", "N", &n, &ndubl, &occ, c, &n, &zero, x, &n);
    for (int i = 1; i <= n; ++i)
      for (int j = 1; j <= i; ++j) p[PK(i, j) - 1] = x[(size_t)(i - 1) * n + (j - 1)];
    free(x);
    return;
  }
  for (int i = 1; i <= n; ++i)
    for (int j = 1; j <= i; ++j) {
      double s = 0.0;
      for (int k = 0; k < ndubl; ++k) s += c[(size_t)k * n + (i - 1)] * c[(size_t)k * n + (j - 1)];
      p[PK(i, j) - 1] = occ * s;
    }
}

/* helect (src/integrals/helect.F90:16-66): E = sum_{i>j} P(H+F) + 0.5 sum_i P(H+F) */
double o_helect(int n, const double* p, const double* h, const double* f) {
  double ed = 0.0, ee = 0.0;
  long long k = 0;
  for (int i = 2; i <= n + 1; ++i) {
    k += 1;
    int jj = i - 1;
    ed += p[k - 1] * (h[k - 1] + f[k - 1]);
    if (i == n + 1) continue;
    for (int q = 1; q <= jj; ++q) ee += p[k + q - 1] * (h[k + q - 1] + f[k + q - 1]);
    k += jj;
  }
  return ee + 0.5 * ed;
}

/* ------------------------------------------------------------------------ */
/* fock2 (src/SCF/fock2.F90:16-374) for molecules (id == 0), with            */
/* jab (src/integrals/jab.F90) and kab (src/integrals/kab.F90) written as     */
/* the index contractions they unroll, and fock1 (src/SCF/fock1.F90).         */
/* ------------------------------------------------------------------------ */
static inline int ifact(int i) { return (i * (i - 1)) / 2; }

/* index of orbital pair (a,b) on an atom (0-based local orbital, a>=b) */
static inline int lpair(int a, int b) {
  if (a < b) { int t = a; a = b; b = t; }
  return a * (a + 1) / 2 + b;
}

void o_fock1(double* f, const double* ptot, const double* pa, const double* w, int ia, int ib) {
  int ilim = ((ib - ia + 1) * (ib - ia + 2)) / 2;
  for (int i = ia; i <= ib; ++i) {
    int iw = i - ia + 1;
    for (int j = ia; j <= i; ++j) {
      int jw = j - ia + 1;
      long long ij = PK(i, j);
      int ijw = (iw * (iw - 1)) / 2 + jw;
      double sum = 0.0;
      for (int k = ia; k <= ib; ++k) {
        int kw = k - ia + 1;
        for (int l = ia; l <= ib; ++l) {
          int lw = l - ia + 1;
          long long ijp = PK(k, l);
          int im = kw > lw ? kw : lw, jm = kw > lw ? lw : kw;
          int klw = (im * (im - 1)) / 2 + jm;
          im = kw > jw ? kw : jw; jm = kw > jw ? jw : kw;
          int ikw = (im * (im - 1)) / 2 + jm;
          im = lw > iw ? lw : iw; jm = lw > iw ? iw : lw;
          int jlw = (im * (im - 1)) / 2 + jm;
          sum += ptot[ijp - 1] * w[(size_t)(klw - 1) * ilim + (ijw - 1)] -
                 pa[ijp - 1] * w[(size_t)(jlw - 1) * ilim + (ikw - 1)];
        }
      }
      f[ij - 1] += sum;
    }
  }
}

/* F += G(P).  nfirst/nlast are 1-based; w is the hcore-ordered integral list. */
void o_fock2(int n, int numat, const int* nfirst, const int* nlast, double* f,
             const double* ptot, const double* p, const double* w) {
  (void)n;
  long long kk = 0;
  for (int ii = 1; ii <= numat; ++ii) {
    int ia = nfirst[ii - 1], ib = nlast[ii - 1];
    int na = ib - ia + 1;
    for (int jj = 1; jj < ii; ++jj) {
      int ja = nfirst[jj - 1], jb = nlast[jj - 1];
      int nb = jb - ja + 1;
      const double* wp = w + kk;
      if (na == 4 && nb == 4) {
        /* Coulomb (jab): F_A(mn) += sum_ls P_B(ls) (mn|ls); F_B(ls) += sum_mn P_A(mn)(mn|ls) */
        double suma[10], sumb[10];
        for (int q = 0; q < 10; ++q) suma[q] = sumb[q] = 0.0;
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) {
            double pA = ptot[PK(ia + a, ia + b) - 1];
            double pB = ptot[PK(ja + a, ja + b) - 1];
            int ab = lpair(a, b);
            for (int q = 0; q < 10; ++q) {
              suma[q] += pA * wp[ab * 10 + q];
              sumb[q] += pB * wp[q * 10 + ab];
            }
          }
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b <= a; ++b) {
            f[PK(ia + a, ia + b) - 1] += sumb[lpair(a, b)];
            f[PK(ja + a, ja + b) - 1] += suma[lpair(a, b)];
          }
        /* Exchange (kab): F(m on A, l on B) -= sum_{n on A, s on B} P(n,s) (mn|ls) */
        for (int m = 0; m < 4; ++m)
          for (int l = 0; l < 4; ++l) {
            double s = 0.0;
            for (int nn = 0; nn < 4; ++nn)
              for (int ss = 0; ss < 4; ++ss)
                s += p[PK(ia + nn, ja + ss) - 1] * wp[lpair(m, nn) * 10 + lpair(l, ss)];
            f[PK(ia + m, ja + l) - 1] -= s;
          }
        kk += 100;
      } else if (na == 4 && nb == 1) {
        /* heavy A, light B (fock2.F90:199-244) */
        long long ll = PK(ja, ja);
        double sumdia = 0.0, sumoff = 0.0;
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b <= a; ++b) {
            int ab = lpair(a, b);
            long long idx = PK(ia + a, ia + b);
            f[idx - 1] += ptot[ll - 1] * wp[ab];
            if (a == b) sumdia += ptot[idx - 1] * wp[ab];
            else sumoff += ptot[idx - 1] * wp[ab];
          }
        f[ll - 1] += sumoff * 2.0 + sumdia;
        for (int m = 0; m < 4; ++m) {
          double s = 0.0;
          for (int nn = 0; nn < 4; ++nn) s += p[PK(ia + nn, ja) - 1] * wp[lpair(m, nn)];
          f[PK(ia + m, ja) - 1] -= s;
        }
        kk += 10;
      } else if (na == 1 && nb == 4) {
        /* light A, heavy B (fock2.F90:245-286) */
        long long ll = PK(ia, ia);
        double sumdia = 0.0, sumoff = 0.0;
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b <= a; ++b) {
            int ab = lpair(a, b);
            long long idx = PK(ja + a, ja + b);
            f[idx - 1] += ptot[ll - 1] * wp[ab];
            if (a == b) sumdia += ptot[idx - 1] * wp[ab];
            else sumoff += ptot[idx - 1] * wp[ab];
          }
        f[ll - 1] += sumoff * 2.0 + sumdia;
        for (int l = 0; l < 4; ++l) {
          double s = 0.0;
          for (int ss = 0; ss < 4; ++ss) s += p[PK(ia, ja + ss) - 1] * wp[lpair(l, ss)];
          f[PK(ia, ja + l) - 1] -= s;
        }
        kk += 10;
      } else {
        /* light - light (fock2.F90:287-297) */
        long long i1 = PK(ia, ia), j1 = PK(ja, ja), ij = PK(ia, ja);
        f[i1 - 1] += ptot[j1 - 1] * wp[0];
        f[j1 - 1] += ptot[i1 - 1] * wp[0];
        f[ij - 1] -= p[ij - 1] * wp[0];
        kk += 1;
      }
    }
    int ilim = (na * (na + 1)) / 2;
    o_fock1(f, ptot, p, w + kk, ia, ib);
    kk += (long long)ilim * ilim;
  }
}

/* The general four-index loop of fock2 (fock2.F90:300-362) with wj = wk = w,
 * restricted to jj < ii: a brute-force second restatement used only to check
 * the fast pair path above. */
void o_fock2_general(int numat, const int* nfirst, const int* nlast, double* f,
                     const double* ptot, const double* p, const double* w) {
  long long kk = 0;
  for (int ii = 1; ii <= numat; ++ii) {
    int ia = nfirst[ii - 1], ib = nlast[ii - 1];
    for (int jj = 1; jj < ii; ++jj) {
      int ja = nfirst[jj - 1], jb = nlast[jj - 1];
      for (int i = ia; i <= ib; ++i) {
        long long ka = ifact(i);
        for (int j = ia; j <= i; ++j) {
          long long kb = ifact(j);
          long long ij = ka + j;
          double aa = (i == j) ? 1.0 : 2.0;
          for (int k = ja; k <= jb; ++k) {
            long long kc = ifact(k);
            long long ik = (i >= k) ? ka + k : 0;
            long long jk = (j >= k) ? kb + k : 0;
            for (int l = ja; l <= k; ++l) {
              long long il = (i >= l) ? ka + l : 0;
              long long jl = (j >= l) ? kb + l : 0;
              long long kl = kc + l;
              double bb = (k == l) ? 1.0 : 2.0;
              kk += 1;
              double aj = w[kk - 1], ak = w[kk - 1];
              if (kl > ij) continue;
              if (i == k && aa + bb < 2.1) {
                f[ij - 1] += aj * ptot[kl - 1];
              } else {
                f[ij - 1] += bb * aj * ptot[kl - 1];
                f[kl - 1] += aa * aj * ptot[ij - 1];
                double a = ak * aa * bb * 0.25;
                if (jl > 0) f[ik - 1] -= a * p[jl - 1];
                if (jk > 0) f[il - 1] -= a * p[jk - 1];
                if (jk > 0) f[jk - 1] -= a * p[il - 1];
                if (jl > 0) f[jl - 1] -= a * p[ik - 1];
              }
            }
          }
        }
      }
    }
    int na = ib - ia + 1, ilim = (na * (na + 1)) / 2;
    o_fock1(f, ptot, p, w + kk, ia, ib);
    kk += (long long)ilim * ilim;
  }
}

/* ------------------------------------------------------------------------ */
/* cnvg (src/SCF/cnvg.F90:16-145), RHF (rhfuhf = 2).                          */
/* ------------------------------------------------------------------------ */
void o_cnvg(int norbs, double* pnew, double* p, double* p1, int niter, double* pl_out) {
  const double rhfuhf = 2.0;
  double pl = 0.0, faca = 0.0, facb = 0.0, fac = 0.0;
  double damp = 1.0e10;
  if (niter > 3) damp = 0.05;
  int extrap = (niter % 3) != 0;
  double sum1 = 0.0;
  long long k = 0;
  for (int i = 1; i <= norbs; ++i) {
    k += i;
    double a = pnew[k - 1];
    sum1 += a;
    double sa = fabs(a - p[k - 1]);
    if (sa > pl) pl = sa;
    if (!extrap) {
      faca += sa * sa;
      double t = a - 2.0 * p[k - 1] + p1[i - 1];
      facb += t * t;
    }
    p1[i - 1] = p[k - 1];
    p[k - 1] = a;
  }
  if (facb > 1.0e-10 && faca < 100.0 * facb) fac = sqrt(faca / facb);
  long long ie = 0;
  double sum2 = 0.0;
  for (int i = 1; i <= norbs; ++i) {
    int ii = i - 1;
    for (int j = 1; j <= ii; ++j) {
      ie += 1;
      double a = pnew[ie - 1];
      p[ie - 1] = a + fac * (a - p[ie - 1]);
      pnew[ie - 1] = p[ie - 1];
    }
    ie += 1;
    if (fabs(p[ie - 1] - p1[i - 1]) > damp) {
      double d = p[ie - 1] - p1[i - 1];
      p[ie - 1] = p1[i - 1] + (d >= 0 ? damp : -damp);
    } else {
      p[ie - 1] = p[ie - 1] + fac * (p[ie - 1] - p1[i - 1]);
    }
    double v = p[ie - 1];
    if (v < 0.0) v = 0.0;
    if (v > rhfuhf) v = rhfuhf;
    p[ie - 1] = v;
    sum2 += v;
    pnew[ie - 1] = v;
  }
  double sum0 = sum1;
  for (;;) {
    double sum = (sum2 > 1.0e-3) ? sum1 / sum2 : 0.0;
    sum1 = sum0;
    if (sum2 < 1.0e-3 || fabs(sum - 1.0) < 1.0e-5) break;
    sum2 = 0.0;
    for (int i = 1; i <= norbs; ++i) {
      long long j = PK(i, i);
      p[j - 1] = p[j - 1] * sum + 1.0e-20;
      if (p[j - 1] < 0.0) p[j - 1] = 0.0;
      if (p[j - 1] > rhfuhf) {
        p[j - 1] = rhfuhf;
        sum1 -= rhfuhf;
      } else {
        sum2 += p[j - 1];
      }
      pnew[j - 1] = p[j - 1];
    }
  }
  *pl_out = pl;
}

/* ------------------------------------------------------------------------ */
/* Pulay DIIS (src/SCF/pulay.F90:16-166) with osinv (src/matrix/osinv.F90,    */
/* full-pivot Gauss-Jordan inverse returning the determinant).                */
/* ------------------------------------------------------------------------ */
typedef struct {
  int mfock, nfock, lfock, start;
  long long linear;
  double* fppf;   /* npulay * mpack  commutators */
  double* fock;   /* npulay * mpack  interleaved (fock(lfock + (i-1)*mfock)) */
  double emat[11 * 11];
} o_pulay_state;

/* Inverse of a general n x n column-major matrix by Gauss-Jordan with full
 * pivoting; d = determinant (0 when singular below tol 1e-8). */
void o_osinv(double* a, int n, double* d_out) {
  const double tol = 1.0e-8;
  int l[16], m[16];
  double d = 1.0;
  for (int k = 0; k < n; ++k) {
    l[k] = k; m[k] = k;
    double biga = a[k * n + k];
    for (int j = k; j < n; ++j)
      for (int i = k; i < n; ++i)
        if (fabs(biga) < fabs(a[j * n + i])) { biga = a[j * n + i]; l[k] = i; m[k] = j; }
    int j = l[k];
    if (j > k)
      for (int i = 0; i < n; ++i) {
        double holo = -a[i * n + k];
        a[i * n + k] = a[i * n + j];
        a[i * n + j] = holo;
      }
    int i = m[k];
    if (i > k)
      for (int jj = 0; jj < n; ++jj) {
        double holo = -a[k * n + jj];
        a[k * n + jj] = a[i * n + jj];
        a[i * n + jj] = holo;
      }
    if (fabs(biga) < tol) { *d_out = 0.0; return; }
    for (int ii = 0; ii < n; ++ii)
      if (ii != k) a[k * n + ii] /= -biga;
    for (int ii = 0; ii < n; ++ii) {
      if (ii == k) continue;
      double holo = a[k * n + ii];
      for (int jj = 0; jj < n; ++jj)
        if (jj != k) a[jj * n + ii] += holo * a[jj * n + k];
    }
    for (int jj = 0; jj < n; ++jj)
      if (jj != k) a[jj * n + k] /= biga;
    d = d * biga < 1.0e10 ? d * biga : 1.0e10;
    a[k * n + k] = 1.0 / biga;
  }
  for (int k = n - 2; k >= 0; --k) {
    int i = l[k];
    if (i > k)
      for (int j = 0; j < n; ++j) {
        double holo = a[k * n + j];
        a[k * n + j] = -a[i * n + j];
        a[i * n + j] = holo;
      }
    int j = m[k];
    if (j > k)
      for (int ii = 0; ii < n; ++ii) {
        double holo = a[ii * n + k];
        a[ii * n + k] = -a[ii * n + j];
        a[ii * n + j] = holo;
      }
  }
  *d_out = d;
}

/* FP - PF of two packed symmetric matrices, packed like the reference's
 * sym_commute + pack_matrix (src/SCF/pulay.F90:168-247). */
static void packed_commutator(int n, const double* p, const double* f, double* out) {
  double* P = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* F = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* X = (double*)calloc((size_t)n * n, sizeof(double));
  for (int i = 1; i <= n; ++i)
    for (int j = 1; j <= n; ++j) {
      P[(size_t)(j - 1) * n + (i - 1)] = p[PK(i, j) - 1];
      F[(size_t)(j - 1) * n + (i - 1)] = f[PK(i, j) - 1];
    }
  if (o_lapack_available() && p_dsymm && n > 64) {
    double one = 1.0, zero = 0.0;
    p_dsymm("L", "U", &n, &n, &one, P, &n, F, &n, &zero, X, &n);
  } else {
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        double fk = F[(size_t)j * n + k];
        for (int i = 0; i < n; ++i) X[(size_t)j * n + i] += P[(size_t)k * n + i] * fk;
      }
  }
  /* mat3 = PF - (PF)^T, packed from the upper triangle (i<=j) */
  for (int j = 1; j <= n; ++j)
    for (int i = 1; i <= j; ++i)
      out[PK(j, i) - 1] = X[(size_t)(j - 1) * n + (i - 1)] - X[(size_t)(i - 1) * n + (j - 1)];
  free(P); free(F); free(X);
}

void o_pulay(int n, int npulay, o_pulay_state* st, double* f, const double* p, double* pl_out) {
  long long linear = (long long)n * (n + 1) / 2;
  if (st->start) {
    st->linear = linear;
    st->mfock = npulay;
    st->nfock = 1;
    st->lfock = 1;
    st->start = 0;
  } else {
    if (st->nfock < st->mfock) st->nfock += 1;
    if (st->lfock != st->mfock) st->lfock += 1;
    else st->lfock = 1;
  }
  int mfock = st->mfock, nfock = st->nfock, lfock = st->lfock;
  long long lbase = (long long)(lfock - 1) * linear;
  for (long long i = 0; i < linear; ++i) st->fock[(size_t)i * mfock + (lfock - 1)] = f[i];
  packed_commutator(n, p, f, st->fppf + lbase);
  int nf1 = nfock + 1;
  int ld = npulay + 1;
#define EM(i, j) st->emat[((j)-1) * ld + ((i)-1)]
  for (int i = 1; i <= nfock; ++i) {
    EM(nf1, i) = -1.0;
    EM(i, nf1) = -1.0;
    double s = 0.0;
    const double* a = st->fppf + (long long)(i - 1) * linear;
    const double* b = st->fppf + lbase;
    for (long long q = 0; q < linear; ++q) s += a[q] * b[q];
    EM(lfock, i) = s;
    EM(i, lfock) = s;
  }
  *pl_out = EM(lfock, lfock) / (double)linear;
  EM(nf1, nf1) = 0.0;
  if (EM(lfock, lfock) < 1.0e-20) return;
  double cst = 1.0 / EM(lfock, lfock);
  for (int j = 1; j <= nfock; ++j)
    for (int i = 1; i <= nfock; ++i) EM(i, j) *= cst;
  double evec[121];
  int l = 0;
  for (int i = 1; i <= nf1; ++i) {
    for (int j = 1; j <= nf1; ++j) evec[l + j - 1] = EM(i, j);
    l += nf1;
  }
  cst = 1.0 / cst;
  for (int j = 1; j <= nfock; ++j)
    for (int i = 1; i <= nfock; ++i) EM(i, j) *= cst;
#undef EM
  double d;
  o_osinv(evec, nf1, &d);
  if (fabs(d) < 1.0e-6) {
    st->start = 1;
    return;
  }
  if (nfock < 2) return;
  int il = nfock * nf1;
  double coeffs[10];
  for (int i = 0; i < nfock; ++i) coeffs[i] = -evec[il + i];
  for (long long i = 0; i < linear; ++i) {
    double s = 0.0;
    long long ii = i * mfock;
    for (int j = 0; j < nfock; ++j) s += coeffs[j] * st->fock[ii + j];
    f[i] = s;
  }
}

/* A sequence of `ncalls` Pulay steps from a fresh start, as iter makes them:
 * call k extrapolates f_seq[k] (mpack) in place using p_seq[k], and pl_seq[k]
 * receives its PL.  Flat arguments so a test can drive it call for call. */
void o_pulay_seq(int n, int npulay, int ncalls, double* f_seq, const double* p_seq,
                 double* pl_seq) {
  long long mpack = (long long)n * (n + 1) / 2;
  o_pulay_state st;
  memset(&st, 0, sizeof(st));
  st.start = 1;
  st.fppf = (double*)calloc((size_t)npulay * mpack, sizeof(double));
  st.fock = (double*)calloc((size_t)npulay * mpack, sizeof(double));
  for (int k = 0; k < ncalls; ++k)
    o_pulay(n, npulay, &st, f_seq + k * mpack, p_seq + k * mpack, pl_seq + k);
  free(st.fppf);
  free(st.fock);
}

/* ------------------------------------------------------------------------ */
/* Pseudo-diagonalization (src/matrix/diag.F90:16-163).                       */
/* vector is n x n column-major (mdim == n); fao packed.                      */
/* Returns the number of rotations applied.                                   */
/* ------------------------------------------------------------------------ */
long long o_diag(const double* fao, double* vector, int nocc, const double* eig, int n) {
  const double bigeps = 1.5e-7;
  int nvirt = n - nocc;
  double* fmo = (double*)malloc(sizeof(double) * (size_t)(nvirt > 0 ? nvirt : 1) * (nocc > 0 ? nocc : 1));
  double* ws = (double*)malloc(sizeof(double) * (size_t)n);
  double tiny = 0.0;
  long long ij = 0;
  for (int i = nocc + 1; i <= n; ++i) {
    long long kk = 0;
    for (int j = 1; j <= n; ++j) {
      double sum = 0.0;
      for (int k = 1; k <= j; ++k) {
        kk += 1;
        sum += fao[kk - 1] * vector[(size_t)(i - 1) * n + (k - 1)];
      }
      if (j != n) {
        long long k2 = kk;
        for (int k = j + 1; k <= n; ++k) {
          k2 += k - 1;
          sum += fao[k2 - 1] * vector[(size_t)(i - 1) * n + (k - 1)];
        }
      }
      ws[j - 1] = sum;
    }
    for (int j = 1; j <= nocc; ++j) {
      ij += 1;
      double sum = 0.0;
      for (int k = 1; k <= n; ++k) sum += ws[k - 1] * vector[(size_t)(j - 1) * n + (k - 1)];
      if (fabs(sum) > tiny) tiny = fabs(sum);
      fmo[ij - 1] = sum;
    }
  }
  tiny *= 0.05;
  ij = 0;
  long long nrot = 0;
  for (int i = nocc + 1; i <= n; ++i)
    for (int j = 1; j <= nocc; ++j) {
      ij += 1;
      if (fabs(fmo[ij - 1]) < tiny) continue;
      double a = eig[j - 1], b = eig[i - 1], c = fmo[ij - 1];
      double d = a - b;
      if (fabs(c / d) < bigeps) continue;
      double e = (d >= 0 ? 1.0 : -1.0) * sqrt(4.0 * c * c + d * d);
      double alpha = sqrt(0.5 * (1.0 + d / e));
      double beta = -(c >= 0 ? 1.0 : -1.0) * sqrt(1.0 - alpha * alpha);
      double* vj = vector + (size_t)(j - 1) * n;
      double* vi = vector + (size_t)(i - 1) * n;
      for (int m = 0; m < n; ++m) {
        double x = vj[m], y = vi[m];
        vj[m] = alpha * x + beta * y;
        vi[m] = alpha * y - beta * x;
      }
      nrot++;
    }
  free(fmo);
  free(ws);
  return nrot;
}

/* ------------------------------------------------------------------------ */
/* iter (src/SCF/iter.F90:16-1107): RHF closed shell, molecule (id == 0),    */
/* no C.I., no Camp-King, fulscf, rand.  Keywords: PULAY, SCFCRT=, ITRY=.      */
/* ------------------------------------------------------------------------ */
typedef struct {
  int norbs, numat, nclose;
  const int* nfirst;
  const int* nlast;
  const double* h;
  const double* w;
  const double* pdiag;   /* NULL: use p_init */
  const double* p_init;  /* optional full packed starting density (total P) */
  double enuclr, atheat;
  int pulay, itrmax, last;
  double scfcrt;
} o_scf_in;

typedef struct {
  double escf, ee;
  int niter, converged;
  double* p;     /* mpack, total density */
  double* f;     /* mpack, final Fock */
  double* c;     /* n*n eigenvectors */
  double* eigs;  /* n */
  double* trace; /* optional: escf per iteration (capacity trace_cap) */
  int trace_cap;
  double* trace_pl;
} o_scf_out;

int o_iter(const o_scf_in* in, o_scf_out* out) {
  const int n = in->norbs;
  const long long mpack = (long long)n * (n + 1) / 2;
  const double enrgy = 23.060547830619029;
  const int npulay = 10;
  double* p = out->p;
  double* f = out->f;
  double* c = out->c;
  double* eigs = out->eigs;
  double* pa = (double*)malloc(sizeof(double) * (size_t)mpack);
  double* pold = (double*)calloc((size_t)mpack, sizeof(double));
  double* pold2 = (double*)calloc((size_t)n, sizeof(double));
  double* fcopy = (double*)malloc(sizeof(double) * (size_t)mpack);
  o_pulay_state pst;
  memset(&pst, 0, sizeof(pst));
  pst.start = 1;
  pst.fppf = (double*)calloc((size_t)npulay * mpack, sizeof(double));
  pst.fock = (double*)calloc((size_t)npulay * mpack, sizeof(double));

  /* initialization (iter.F90:91-251) */
  double ten = 10.0, tenold = 10.0, bshift = -80.0, shift = 1.0, shfto = 0.0, shftbo = 0.0;
  double shiftb = 0.0, shfmax = 20.0, plchek = 0.005, pl = 1.0, trans = 0.2;
  int itrmax = in->itrmax > 0 ? in->itrmax : 2000;
  int okpuly = in->pulay;
  int oknewd = fabs(bshift) < 0.001;
  int newdg = 0;
  if (fabs(bshift) > 1e-20) ten = bshift;
  int na1el = in->nclose, nb1el = in->nclose, na2el = in->nclose;
  double w1 = na1el / (na1el + 1.0e-6 + nb1el);
  if (w1 < 1.0e-6) w1 = 0.5;
  memset(p, 0, sizeof(double) * (size_t)mpack);
  memset(pa, 0, sizeof(double) * (size_t)mpack);
  if (in->p_init) {
    memcpy(p, in->p_init, sizeof(double) * (size_t)mpack);
    for (long long q = 0; q < mpack; ++q) pa[q] = p[q] * 0.5;
  } else {
    for (int i = 1; i <= n; ++i) {
      long long j = PK(i, i);
      p[j - 1] = in->pdiag[i - 1];
      pa[j - 1] = p[j - 1] * w1;
    }
  }
  memcpy(pold, pa, sizeof(double) * (size_t)mpack);
  if (in->p_init) memcpy(pold, p, sizeof(double) * (size_t)mpack);
  for (int i = 1; i <= n; ++i) pold2[i - 1] = pold[PK(i, i) - 1];
  double scfcrt = in->scfcrt > 0 ? in->scfcrt : 1.0e-4;
  if (scfcrt < 1e-12) scfcrt = 1e-12;
  double selcon = scfcrt;
  /* gs: selcon = min(min(scfcrt*100,0.1), max(scfcrt*gnorm^3, scfcrt)) with gnorm = 0 */
  {
    double a = scfcrt * 100.0 < 0.1 ? scfcrt * 100.0 : 0.1;
    double b = scfcrt;
    selcon = a < b ? a : b;
  }
  double pltest = 0.05 * sqrt(fabs(selcon));
  if (in->last == 1) newdg = 0;
  int iredy = 1, niter = 0, ready = 0;
  double eold = 100.0, diff = 0.0, sellim = 0.0, escf = 0.0, ee = 0.0;
  int ihomo = in->nclose > 1 ? in->nclose : 1;
  int converged = 0;
  for (int i = 0; i < n; ++i) eigs[i] = 0.0;
  for (;;) {
    niter += 1;
    /* Fock matrix with level shift (iter.F90:378-465) */
    if (fabs(shift) > 1e-10 && bshift != 0.0) {
      if (niter > 1) {
        if (newdg) {
          if (diff > 0.0) {
            shift = 1.0;
            if (diff > 1) newdg = 0;
          } else {
            shift = -0.1;
          }
        } else {
          if (ihomo < n) shift = ten + eigs[ihomo] - eigs[ihomo - 1] + shift;
          else shift = 0.0;
        }
        if (diff > 0.0) {
          if (shift > 4.0) shfmax = 4.5;
          if (shift > shfmax) shfmax = (shfmax - 0.5 > 0.0) ? shfmax - 0.5 : 0.0;
        }
        shift = shift < shfmax ? shift : shfmax;
        shift = shift > -20.0 ? shift : -20.0;
        if (fabs(shift - shfmax) < 1e-5) shfmax += 0.01;
        if (okpuly || fabs(bshift - 4.44) < 1e-5) {
          shift = -8.0;
          if (newdg) shift = 0.0;
        }
      }
      tenold = ten;
      if (pl > plchek) {
        shftbo = shiftb;
        shfto = shift;
      } else {
        shiftb = shftbo;
        shift = shfto;
      }
      for (int i = ihomo; i < n; ++i) eigs[i] += shift;
      for (long long i = 1; i <= mpack; ++i) f[i - 1] = in->h[i - 1] + shift * pa[i - 1] + 1.0e-16 * (double)i;
      for (int i = 1; i <= n; ++i) f[PK(i, i) - 1] -= shift;
    } else if (in->last == 0 && niter < 2) {
      double random = 0.001;
      for (long long i = 0; i < mpack; ++i) {
        random = -random;
        f[i] = in->h[i] + random;
      }
    } else {
      memcpy(f, in->h, sizeof(double) * (size_t)mpack);
    }
  label320:
    o_fock2(n, in->numat, in->nfirst, in->nlast, f, p, pa, in->w);
    if (niter >= itrmax) break;
    ee = 2.0 * o_helect(n, pa, in->h, f);
    escf = (ee + in->enuclr) * enrgy + in->atheat;
    diff = escf - eold;
    if (diff > 0) ten = ten - 1.0;
    else ten = ten * 0.975 + 0.05;
    sellim = selcon > 1e-15 * (fabs(ee) > 1.0 ? fabs(ee) : 1.0) ? selcon : 1e-15 * (fabs(ee) > 1.0 ? fabs(ee) : 1.0);
    if (out->trace && niter <= out->trace_cap) {
      out->trace[niter - 1] = escf;
      if (out->trace_pl) out->trace_pl[niter - 1] = pl;
    }
    if (niter > 4 && (pl == 0.0 || (pl < pltest && fabs(diff) < sellim)) && ready) {
      if (fabs(shift) < 1e-10) { converged = 1; break; }
      shift = 0.0;
      shiftb = 0.0;
      memcpy(f, in->h, sizeof(double) * (size_t)mpack);
      goto label320;
    }
    ready = iredy > 0 && (fabs(diff) < sellim * 10.0 || pl == 0.0);
    iredy += 1;
    eold = escf;
    if (newdg && okpuly && iredy > 1) o_pulay(n, npulay, &pst, f, pa, &pl);
    o_eigenvectors(n, c, f, eigs);
    o_density(n, na2el, 2.0, c, p);
    if (!(newdg && okpuly)) o_cnvg(n, p, pold, pold2, niter, &pl);
    for (long long q = 0; q < mpack; ++q) pa[q] = p[q] * 0.5;
    oknewd = pl < sellim || oknewd;
    newdg = (pl < trans && oknewd) || newdg;
    if (pl < trans * 0.3333) oknewd = 1;
  }
  /* exit (iter.F90:1052-1106) */
  ee = 2.0 * o_helect(n, pa, in->h, f);
  memcpy(fcopy, f, sizeof(double) * (size_t)mpack);
  o_eigenvectors(n, c, fcopy, eigs);
  if (in->last == 1) o_phase_lock(n, c);
  escf = (ee + in->enuclr) * enrgy + in->atheat;
  out->escf = escf;
  out->ee = ee;
  out->niter = niter;
  out->converged = converged;
  free(pa); free(pold); free(pold2); free(fcopy);
  free(pst.fppf); free(pst.fock);
  return converged ? 0 : 1;
}

/* Flat-argument wrapper for ctypes callers. */
int o_iter_flat(int norbs, int numat, int nclose, const int* nfirst, const int* nlast,
                const double* h, const double* w, const double* pdiag, const double* p_init,
                double enuclr, double atheat, int pulay, int itrmax, double scfcrt, int last,
                double* p, double* f, double* c, double* eigs, double* trace, double* trace_pl,
                int trace_cap, double* escf, double* ee, int* niter) {
  o_scf_in in;
  memset(&in, 0, sizeof(in));
  in.norbs = norbs; in.numat = numat; in.nclose = nclose;
  in.nfirst = nfirst; in.nlast = nlast; in.h = h; in.w = w; in.pdiag = pdiag; in.p_init = p_init;
  in.enuclr = enuclr; in.atheat = atheat; in.pulay = pulay; in.itrmax = itrmax;
  in.scfcrt = scfcrt; in.last = last;
  o_scf_out out;
  memset(&out, 0, sizeof(out));
  out.p = p; out.f = f; out.c = c; out.eigs = eigs;
  out.trace = trace; out.trace_pl = trace_pl; out.trace_cap = trace_cap;
  int rc = o_iter(&in, &out);
  *escf = out.escf;
  *ee = out.ee;
  *niter = out.niter;
  return rc;
}
