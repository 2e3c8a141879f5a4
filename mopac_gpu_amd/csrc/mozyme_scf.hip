// MOZYME's sparse Fock build on the device: buildf(f, partf, 0) with
// fock2z/fz2n (src/MOZYME/buildf.F90:16-52, src/MOZYME/fock2z.F90:16-1362)
// for a discrete system (ione = 1) with stored near-pair integrals and
// point-charge terms for the other pairs evaluated from the parameters (the
// SEMIDIRECT storage, direct = false, semidr = true; DIRECT gives the same
// numbers with the near-pair integrals recomputed).
//
// Layout: P, F and H live in the MOZYME sparse storage (atom-pair blocks at
// nijbo, a > b as an iorbs(a) x iorbs(b) block with b's orbital fastest,
// a == b as the packed lower triangle).  The near-pair part of fz2n is the
// conventional fock2 on those blocks (jab/kab/fock1 with the total density),
// so it runs as
//   1. scatter the sparse P into a dense packed P (and P/2), far blocks 0,
//   2. the conventional two-pass Fock kernels over an integral array in hcore
//      order whose far-pair blocks are zero,
//   3. gather F = H + G back into the sparse blocks,
// then one workgroup per atom adds the monopole (and, for the nijbo = -2
// shell, the dipole) terms of every far pair to that atom's own diagonal
// block, so no two workgroups write the same element.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <vector>

#include "ctx.h"

namespace mhip {

struct MzFock {
  int numat = 0, norbs = 0;
  long long mpack = 0;  // sparse storage length
  DevBuf map;           // sparse element -> dense packed index (-1: slack between blocks)
  DevBuf hs;            // sparse H
  DevBuf pd, pad, fd;   // dense packed P, P/2, F
  DevBuf nijbo;         // numat x numat (column-major, symmetric)
  DevBuf iorbs, diag;   // orbital count and sparse diagonal-block offset per atom
  DevBuf xyz, am, dd, ad;
  DevBuf qe;            // electron population per atom
  DevBuf ps, fs;        // sparse P and F (device copies for the host entry point)
  int feather = 0;      // PM7's smooth switch to point charges
  double trunc1 = 0.0, trunc2 = 0.0;
};

namespace {

constexpr double kA0 = 0.529177210903;   // Bohr radius, Angstrom
constexpr double kEV = 27.211386245988;  // Hartree, eV

__global__ void k_mzf_scatter(long long m, const int* __restrict__ map, const double* __restrict__ ps,
                              double* __restrict__ pd, double* __restrict__ pad) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < m;
       e += (long long)gridDim.x * blockDim.x) {
    const int d = map[e];
    if (d < 0) continue;  // slack between blocks
    const double v = ps[e];
    pd[d] = v;
    pad[d] = 0.5 * v;
  }
}

__global__ void k_mzf_gather(long long m, const int* __restrict__ map, const double* __restrict__ hs,
                             const double* __restrict__ fd, double* __restrict__ fs) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < m;
       e += (long long)gridDim.x * blockDim.x)
  {
    const int d = map[e];
    fs[e] = hs[e] + (d >= 0 ? fd[d] : 0.0);
  }
}

// chrge_for_MOZYME (src/MOZYME/chrge_for_MOZYME.F90:16-39): electrons per atom
__global__ void k_mzf_qe(int numat, const int* __restrict__ iorbs, const int* __restrict__ diag,
                         const double* __restrict__ ps, double* __restrict__ qe) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= numat) return;
  long long o = diag[a];
  double sum = 0.0;
  for (int j = 1; j <= iorbs[a]; ++j) {
    o += j;
    sum += ps[o - 1];
  }
  qe[a] = sum;
}

// The far-pair terms of fz2n (fock2z.F90:1074-1200) for atom A (with
// l_feather, PM7's default, the monopole blends into the bare Coulomb
// ev a0 / r and the dipole integrals fade out below trunc1): over every
// atom B whose pair is not stored, the monopole qe(B) w1 on A's diagonal;
// for the dipole shell (nijbo = -2) also qe(B) times A's dipole integrals on
// A's s-p elements (A with p orbitals) and 2 P_B(s,p) times B's dipole
// integrals on A's diagonal (B with p orbitals).  The dipole integral of atom
// X seen from Y is ev (1/(2 rp) - 1/(2 rm)) along the unit vector from Y to X,
// rp/m = sqrt((r/a0 +- dd(X))^2 + (1/(2 ad(X)) + 1/(2 am(Y)))^2), which is
// the reference's w5..w7 for X = ii and w2..w4 for X = jj.
__global__ __launch_bounds__(256) void k_mzf_far(int numat, const int* __restrict__ nijbo,
                                                 const int* __restrict__ iorbs, const int* __restrict__ diag,
                                                 const double* __restrict__ xyz, const double* __restrict__ am,
                                                 const double* __restrict__ dd, const double* __restrict__ ad,
                                                 const double* __restrict__ qe, const double* __restrict__ ps,
                                                 int feather, double trunc1, double trunc2,
                                                 double* __restrict__ fs) {
  const int A = blockIdx.x;
  const int na = iorbs[A];
  double mono = 0.0, dg = 0.0, sp0 = 0.0, sp1 = 0.0, sp2 = 0.0;
  if (na > 0) {
    const double xa = xyz[3 * A], ya = xyz[3 * A + 1], za = xyz[3 * A + 2];
    for (int B = threadIdx.x; B < numat; B += 256) {
      if (B == A) continue;
      const int nb = iorbs[B];
      if (nb == 0) continue;
      const int o = nijbo[B + (size_t)A * numat];
      if (o >= 0) continue;
      const double dx = xa - xyz[3 * B], dy = ya - xyz[3 * B + 1], dz = za - xyz[3 * B + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double aee = 0.5 / am[A] + 0.5 / am[B];
      double w1 = kEV / sqrt(r2 / (kA0 * kA0) + aee * aee);
      const double r = sqrt(r2);
      // PM7's smooth switch to point charges (to_point, src/integrals/mndod.F90:3462-3486)
      double cf = 1.0;
      if (feather) {
        cf = r < trunc1 ? 1.0 - exp(-(r - trunc1) * (r - trunc1) * trunc2) : 0.0;
        w1 = w1 * cf + (1.0 - cf) * (kEV * kA0 / r);
      }
      mono += qe[B] * w1;
      if (o != -2) continue;
      double ux = dx / r, uy = dy / r, uz = dz / r;  // from B to A
      if (fabs(uz) > 0.99999999) uz = copysign(1.0, uz);
      if (na > 1) {
        const double da = dd[A], ade = 0.5 / ad[A] + 0.5 / am[B];
        const double rp = sqrt((r / kA0 + da) * (r / kA0 + da) + ade * ade);
        const double rm = sqrt((r / kA0 - da) * (r / kA0 - da) + ade * ade);
        double s = kEV * (0.5 / rp - 0.5 / rm);
        if (feather) s = s * cf;
        sp0 += qe[B] * (s * ux);
        sp1 += qe[B] * (s * uy);
        sp2 += qe[B] * (s * uz);
      }
      if (nb > 1) {
        const double da = dd[B], ade = 0.5 / ad[B] + 0.5 / am[A];
        const double rp = sqrt((r / kA0 + da) * (r / kA0 + da) + ade * ade);
        const double rm = sqrt((r / kA0 - da) * (r / kA0 - da) + ade * ade);
        double s = kEV * (0.5 / rp - 0.5 / rm);  // along -u
        if (feather) s = s * cf;
        const long long j1 = diag[B];
        dg += ps[j1 + 1] * (s * -ux) * 2.0 + ps[j1 + 3] * (s * -uy) * 2.0 + ps[j1 + 6] * (s * -uz) * 2.0;
      }
    }
  }
  __shared__ double red[5][4];
  double v[5] = {mono, dg, sp0, sp1, sp2};
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) red[k][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0 && na > 0) {
    double t[5];
    for (int k = 0; k < 5; ++k) t[k] = ((red[k][0] + red[k][1]) + red[k][2]) + red[k][3];
    const long long i1 = diag[A];
    long long o = i1;
    for (int i = 1; i <= na; ++i) {
      o += i;
      fs[o - 1] += t[0] + t[1];
    }
    if (na > 1) {
      fs[i1 + 1] += t[2];
      fs[i1 + 3] += t[3];
      fs[i1 + 6] += t[4];
    }
  }
}

inline int nblk(long long n) { return (int)std::min<long long>(4096, (n + 255) / 256); }

}  // namespace

void mz_release(MzFock* z) {
  if (!z) return;
  for (DevBuf* b : {&z->map, &z->hs, &z->pd, &z->pad, &z->fd, &z->nijbo, &z->iorbs, &z->diag, &z->xyz, &z->am,
                    &z->dd, &z->ad, &z->qe, &z->ps, &z->fs})
    b->release();
  delete z;
}

// f (device, sparse) = h + G(p) for p (device, sparse)
void mz_fock2z(mhip_ctx* c, const double* ps, double* fs) {
  MzFock* z = c->mz;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_mzf_scatter, dim3(nblk(z->mpack)), dim3(256), 0, s, z->mpack, z->map.as<int>(), ps,
                     z->pd.as<double>(), z->pad.as<double>());
  MHIP_LAUNCH_CHECK();
  FockArgs a = fock_args(c, kInitH, 0.0);
  a.p = z->pd.as<double>();
  a.pa = z->pad.as<double>();
  a.f = z->fd.as<double>();
  a.fin = z->fd.as<double>();
  fock_build(s, a);
  hipLaunchKernelGGL(k_mzf_gather, dim3(nblk(z->mpack)), dim3(256), 0, s, z->mpack, z->map.as<int>(),
                     z->hs.as<double>(), z->fd.as<double>(), fs);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mzf_qe, dim3((z->numat + 255) / 256), dim3(256), 0, s, z->numat, z->iorbs.as<int>(),
                     z->diag.as<int>(), ps, z->qe.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mzf_far, dim3(z->numat), dim3(256), 0, s, z->numat, z->nijbo.as<int>(), z->iorbs.as<int>(),
                     z->diag.as<int>(), z->xyz.as<double>(), z->am.as<double>(), z->dd.as<double>(),
                     z->ad.as<double>(), z->qe.as<double>(), ps, z->feather, z->trunc1, z->trunc2, fs);
  MHIP_LAUNCH_CHECK();
}

}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_mozyme_setup(mhip_ctx* c, const mhip_mozyme_system* sys) {
  try {
    MHIP_REQUIRE(c && sys, "mhip_mozyme_setup: null argument");
    MHIP_CHECK(hipSetDevice(c->device));
    const int n = sys->numat;
    MHIP_REQUIRE(n > 0 && sys->nfirst && sys->nlast && sys->nijbo && sys->xyz && sys->h && sys->w && sys->am &&
                     sys->dd && sys->ad,
                 "mhip_mozyme_setup: missing array");
    std::vector<int> no(n), nf0(n);
    for (int a = 0; a < n; ++a) {
      nf0[a] = sys->nfirst[a] - 1;
      no[a] = sys->nlast[a] - sys->nfirst[a] + 1;
      MHIP_REQUIRE(no[a] == 1 || no[a] == 4, "mhip_mozyme_setup: only s (H) and sp (C, N, O) atoms");
    }
    const int norbs = nf0[n - 1] + no[n - 1];
    MHIP_REQUIRE(norbs == sys->norbs, "mhip_mozyme_setup: nlast(numat) must equal norbs");
    // sparse -> dense map, checking every stored block lies inside mpack
    std::vector<int> map((size_t)sys->mpack, -1), diag(n);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b <= a; ++b) {
        const int o = sys->nijbo[a + (size_t)b * n];
        MHIP_REQUIRE(o == sys->nijbo[b + (size_t)a * n], "mhip_mozyme_setup: nijbo must be symmetric");
        if (a == b) MHIP_REQUIRE(o >= 0, "mhip_mozyme_setup: every diagonal block must be stored");
        if (o < 0) {
          MHIP_REQUIRE(o == -1 || o == -2, "mhip_mozyme_setup: nijbo entries are offsets, -1 or -2");
          continue;
        }
        const int len = (a == b) ? no[a] * (no[a] + 1) / 2 : no[a] * no[b];
        MHIP_REQUIRE((long long)o + len <= sys->mpack, "mhip_mozyme_setup: a block lies beyond mpack");
        if (a == b) diag[a] = o;
        int k = 0;
        for (int x = 0; x < no[a]; ++x)
          for (int y = 0; y < (a == b ? x + 1 : no[b]); ++y, ++k) {
            const long long r = nf0[a] + x, cc = nf0[b] + y;
            MHIP_REQUIRE(map[(size_t)o + k] < 0, "mhip_mozyme_setup: overlapping blocks in nijbo");
            map[(size_t)o + k] = (int)(r * (r + 1) / 2 + cc);
          }
      }
    // integrals in hcore order for every pair, far-pair blocks zero; the
    // MOZYME array holds the near pairs and the one-centre blocks in that order
    std::vector<double> wconv;
    long long kr = 0;
    for (int a = 0; a < n; ++a) {
      const int la = no[a] * (no[a] + 1) / 2;
      for (int b = 0; b < a; ++b) {
        const int lb = no[b] * (no[b] + 1) / 2;
        const size_t len = (size_t)la * lb;
        if (sys->nijbo[a + (size_t)b * n] >= 0) {
          MHIP_REQUIRE(kr + (long long)len <= sys->nw, "mhip_mozyme_setup: integral array too short");
          wconv.insert(wconv.end(), sys->w + kr, sys->w + kr + len);
          kr += (long long)len;
        } else {
          wconv.insert(wconv.end(), len, 0.0);
        }
      }
      const size_t len = (size_t)la * la;
      MHIP_REQUIRE(kr + (long long)len <= sys->nw, "mhip_mozyme_setup: integral array too short");
      wconv.insert(wconv.end(), sys->w + kr, sys->w + kr + len);
      kr += (long long)len;
    }
    MHIP_REQUIRE(kr == sys->nw, "mhip_mozyme_setup: integral array longer than the stored pairs need");
    std::vector<double> hzero((size_t)mpack_of(norbs), 0.0);
    build_system(c, n, sys->nfirst, sys->nlast, norbs, (long long)wconv.size(), hzero.data(), wconv.data());
    if (!c->mz) c->mz = new MzFock();
    MzFock* z = c->mz;
    z->numat = n;
    z->norbs = norbs;
    z->feather = sys->feather;
    z->trunc1 = sys->trunc1;
    z->trunc2 = sys->trunc2;
    z->mpack = sys->mpack;
    hipStream_t s = c->stream;
    upload(z->map, map.data(), sizeof(int) * map.size(), s);
    upload(z->hs, sys->h, sizeof(double) * sys->mpack, s);
    upload(z->nijbo, sys->nijbo, sizeof(int) * (size_t)n * n, s);
    upload(z->iorbs, no.data(), sizeof(int) * n, s);
    upload(z->diag, diag.data(), sizeof(int) * n, s);
    upload(z->xyz, sys->xyz, sizeof(double) * 3 * n, s);
    upload(z->am, sys->am, sizeof(double) * n, s);
    upload(z->dd, sys->dd, sizeof(double) * n, s);
    upload(z->ad, sys->ad, sizeof(double) * n, s);
    z->qe.ensure(sizeof(double) * n);
    const size_t dense = sizeof(double) * (size_t)mpack_of(norbs);
    for (DevBuf* b : {&z->pd, &z->pad, &z->fd}) {
      b->ensure(dense);
      MHIP_CHECK(hipMemsetAsync(b->ptr, 0, dense, s));
    }
    z->ps.ensure(sizeof(double) * sys->mpack);
    z->fs.ensure(sizeof(double) * sys->mpack);
    MHIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    return 1;
  }
}

int mhip_mozyme_fock2z(mhip_ctx* c, const double* p, double* f) {
  try {
    MHIP_REQUIRE(c && c->mz && p && f, "mhip_mozyme_fock2z: no MOZYME system set up");
    MHIP_CHECK(hipSetDevice(c->device));
    MzFock* z = c->mz;
    hipStream_t s = c->stream;
    MHIP_CHECK(hipMemcpyAsync(z->ps.ptr, p, sizeof(double) * z->mpack, hipMemcpyHostToDevice, s));
    mz_fock2z(c, z->ps.as<double>(), z->fs.as<double>());
    MHIP_CHECK(hipMemcpyAsync(f, z->fs.ptr, sizeof(double) * z->mpack, hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    return 1;
  }
}

}  // extern "C"

// ---- the MOZYME SCF loop (iter_for_MOZYME) ---------------------------------

namespace {

// LMO arrays the loop owns (tidy moves LMOs inside them)
struct LmoStore {
  int nmo = 0;
  std::vector<int> nc, nnc, ncmo, ic;
  std::vector<double> c;
};

// tidy (src/MOZYME/tidy.F90:16-244) for a single point: drop every atom whose
// coefficients in an LMO have a squared sum <= thresh, compress, then spread
// the free space over the LMOs from the last one down (LMO 1 stays at 0).
// Returns false when the storage is too small for the next annihilation
// step (the reference then re-allocates 60% larger arrays).
bool tidy(LmoStore& L, const std::vector<int>& iorbs, int numat, int norbs, double thresh) {
  const int nmo = L.nmo;
  if (nmo == 0) return true;
  const int n01 = (int)L.ic.size(), n02 = (int)L.c.size();
  std::vector<int> iused(nmo + 1, 0);
  int ln = 0, mn = 0, in = 0;
  for (int i = 0; i < nmo; ++i) {
    int ll = L.nnc[i];
    int mm = L.ncmo[i];
    L.ncmo[i] = mn;
    L.nnc[i] = (i == 0) ? 0 : L.nnc[i - 1] + in;
    in = 0;
    for (int j1 = 0; j1 < L.nc[i]; ++j1) {
      const int atom = L.ic[ll++];
      const int no = iorbs[atom - 1];
      double sum = 0.0;
      for (int k = 0; k < no; ++k) sum += L.c[mm + k] * L.c[mm + k];
      if (sum > thresh) {
        for (int k = 0; k < no; ++k) L.c[mn + k] = L.c[mm + k];
        mn += no;
        ++in;
        L.ic[ln++] = atom;
      }
      mm += no;
    }
    L.nc[i] = in;
    iused[i + 1] = mn;
  }
  int ispace = (n01 - ln) / nmo, jspace = (n02 - mn) / nmo;
  const int per = n01 / nmo;
  if (per + ispace < numat && ispace < std::max(per / 5, 20)) return false;
  if (ispace > numat && jspace > nmo) ispace = jspace = 0;
  int jtop = n01, mtop = n02;
  int jsav = n01 - ln - ispace * nmo, msav = n02 - mn - jspace * nmo;
  for (int i = nmo - 1; i >= 1; --i) {
    int j = std::min(L.nc[i] + ispace, numat);
    if (j < L.nc[i] + ispace) {
      jsav += L.nc[i] + ispace - j;
      jtop -= j;
    } else {
      const int jdash = std::min(numat, std::max(jsav, 0) + j);
      jtop -= jdash;
      jsav = jsav - jdash + j;
    }
    const int l = L.nnc[i];
    // the moved LMO must land at or above its compressed position (and so
    // never over an LMO below it); the reference's bookkeeping keeps it so
    if (jtop < l) return false;
    L.nnc[i] = jtop;
    for (int k = L.nc[i] - 1; k >= 0; --k) L.ic[jtop + k] = L.ic[l + k];
    const int n = iused[i + 1] - iused[i];
    j = std::min(n + jspace, norbs);
    if (j < n + jspace) {
      msav += n + jspace - j;
      mtop -= j;
    } else {
      const int mdash = std::min(norbs, std::max(msav, 0) + j);
      mtop -= mdash;
      msav = msav - mdash + j;
    }
    if (mtop < iused[i]) return false;
    L.ncmo[i] = mtop;
    for (int k = n - 1; k >= 0; --k) L.c[mtop + k] = L.c[iused[i] + k];
  }
  return true;
}

// check (src/MOZYME/check.F90:16-113): renormalize every LMO; false when the
// summed normalization error exceeds 0.1 (the reference stops the job)
bool check(LmoStore& L, const std::vector<int>& iorbs) {
  double error = 0.0;
  std::vector<double> ws(L.nmo);
  for (int i = 0; i < L.nmo; ++i) {
    int n = 0;
    double sum = 0.0;
    for (int j = 0; j < L.nc[i]; ++j) {
      const int atom = L.ic[L.nnc[i] + j];
      for (int mm = 0; mm < iorbs[atom - 1]; ++mm, ++n) sum += L.c[L.ncmo[i] + n] * L.c[L.ncmo[i] + n];
    }
    error += std::fabs(1.0 - sum);
    ws[i] = sum;
  }
  for (int i = 0; i < L.nmo; ++i) {
    const double sc = 1.0 / std::sqrt(ws[i]);
    int n = 0;
    for (int j = 0; j < L.nc[i]; ++j) {
      const int atom = L.ic[L.nnc[i] + j];
      for (int mm = 0; mm < iorbs[atom - 1]; ++mm, ++n) L.c[L.ncmo[i] + n] *= sc;
    }
  }
  return error <= 0.1;
}

}  // namespace

extern "C" {

int mhip_mozyme_scf(mhip_ctx* c, mhip_lmo_set* lmo, const mhip_mozyme_opts* o, mhip_mozyme_result* res,
                    double* p_out, double* f_out) {
  try {
    MHIP_REQUIRE(c && c->mz && lmo && o && res, "mhip_mozyme_scf: set up the system first (mhip_mozyme_setup)");
    MHIP_CHECK(hipSetDevice(c->device));
    MzFock* z = c->mz;
    const int numat = z->numat, norbs = z->norbs;
    const long long mp = z->mpack;
    MHIP_REQUIRE(lmo->numat == numat && lmo->norbs == norbs && lmo->mpack == mp,
                 "mhip_mozyme_scf: LMO set and system differ");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<int> iorbs(numat), nijbo((size_t)numat * numat);
    for (int a = 0; a < numat; ++a) iorbs[a] = lmo->nlast[a] - lmo->nfirst[a] + 1;
    std::copy(lmo->nijbo, lmo->nijbo + (size_t)numat * numat, nijbo.begin());
    // working copies of the LMOs
    LmoStore occ, vir;
    occ.nmo = lmo->nocc;
    vir.nmo = lmo->nvir;
    occ.nc.assign(lmo->ncf, lmo->ncf + lmo->nocc);
    occ.nnc.assign(lmo->nncf, lmo->nncf + lmo->nocc);
    occ.ncmo.assign(lmo->ncocc, lmo->ncocc + lmo->nocc);
    occ.ic.assign(lmo->icocc, lmo->icocc + lmo->icocc_dim);
    occ.c.assign(lmo->cocc, lmo->cocc + lmo->cocc_dim);
    vir.nc.assign(lmo->nce, lmo->nce + lmo->nvir);
    vir.nnc.assign(lmo->nnce, lmo->nnce + lmo->nvir);
    vir.ncmo.assign(lmo->ncvir, lmo->ncvir + lmo->nvir);
    vir.ic.assign(lmo->icvir, lmo->icvir + lmo->icvir_dim);
    vir.c.assign(lmo->cvir, lmo->cvir + lmo->cvir_dim);
    mhip_lmo_set L = *lmo;
    auto bind = [&] {
      L.ncf = occ.nc.data();
      L.nncf = occ.nnc.data();
      L.ncocc = occ.ncmo.data();
      L.icocc = occ.ic.data();
      L.cocc = occ.c.data();
      L.nce = vir.nc.data();
      L.nnce = vir.nnc.data();
      L.ncvir = vir.ncmo.data();
      L.icvir = vir.ic.data();
      L.cvir = vir.c.data();
    };
    bind();
    // idiag: 1-based position of P(i, i) in the sparse storage
    std::vector<int> idiag(norbs), diag(numat);
    for (int a = 0, l = 0; a < numat; ++a) {
      int j = nijbo[a + (size_t)a * numat];
      for (int k = 1; k <= iorbs[a]; ++k) {
        j += k;
        idiag[l++] = j;
      }
    }
    // stored off-diagonal blocks (offset, size), in eimp's order (i, then j < i)
    std::vector<std::pair<int, int>> offdiag;
    for (int i = 0; i < numat; ++i)
      for (int j = 0; j < i; ++j) {
        const int k = nijbo[i + (size_t)j * numat];
        if (k >= 0) offdiag.emplace_back(k, iorbs[i] * iorbs[j]);
      }
    std::vector<double> h((size_t)mp), P((size_t)mp), F((size_t)mp), pold((size_t)mp, 0.0), p1(norbs, 0.0),
        p2(norbs, 0.0), p3(norbs, 0.0);
    MHIP_CHECK(hipMemcpy(h.data(), z->hs.ptr, sizeof(double) * mp, hipMemcpyDeviceToHost));
    std::vector<double> eigs(std::max(lmo->nocc, 1), 0.0), eigv(std::max(lmo->nvir, 1), 0.0);
    const int fmo_dim = o->fmo_dim > 0 ? o->fmo_dim : (int)std::min<long long>(
        (long long)lmo->nocc * lmo->nvir, 50LL * std::max(lmo->nocc, lmo->nvir) * 64);
    std::vector<double> fmo(std::max(fmo_dim, 1));
    std::vector<int> ifmo(2 * (size_t)std::max(fmo_dim, 1)), nfmo(std::max(lmo->nvir, 1), 0);
    mhip_diagg1_state st1{};
    st1.reset = 1;
    st1.nfmo = nfmo.data();
    mhip_diagg2_state st2{};
    st2.reset = 1;
    st2.damp = o->damp > 0 ? o->damp : 1.0;
    st2.thresh = o->thresh > 0 ? o->thresh : 1.0e-13;
    st2.order_free = o->order_free ? 1 : 0;
    const double selcon = o->scfcrt > 0 ? o->scfcrt : 1.0e-2;  // scfcri (src/MOZYME/scfcri.F90)
    const int itrmax = o->itrmax > 0 ? o->itrmax : 2000;
    const double fpc9 = mhip_ev_to_kcal();
    auto check_rc = [](int rc) {
      if (rc != 0) throw HipError{get_error()};
    };
    auto energy = [&](double& ee) {
      check_rc(mhip_mozyme_helecz(c, numat, lmo->nfirst, lmo->nlast, nijbo.data(), mp, P.data(), h.data(),
                                  F.data(), &ee));
      return (ee + o->enuclr) * fpc9 + o->atheat;
    };
    // iter_for_MOZYME.F90:433-478: density from the LMOs, Fock, energy
    check_rc(mhip_mozyme_density(c, &L, 0, nullptr, P.data()));
    check_rc(mhip_mozyme_fock2z(c, P.data(), F.data()));
    double ee = 0.0;
    double escf = energy(ee);
    double eold = 0.0, ediff = 0.0, shift = 0.0;
    int niter = 0, idiagg = 0, lstart = 0;
    bool three = true, scf1 = false, okscf = false;
    for (;;) {
      MHIP_REQUIRE(tidy(occ, iorbs, numat, norbs, st2.thresh) && tidy(vir, iorbs, numat, norbs, st2.thresh),
                   "mhip_mozyme_scf: LMO storage too small for the next step (enlarge icocc/cocc/icvir/cvir)");
      MHIP_REQUIRE(check(occ, iorbs) && check(vir, iorbs), "mhip_mozyme_scf: LMO normalization error > 0.1");
      // eimp (src/MOZYME/eimp.F90:16-45): the s-s slot of each stored
      // off-diagonal block of P holds the sum of squares of that F block
      for (const auto& b : offdiag) {
        double sum = 0.0;
        for (int m = 0; m < b.second; ++m) sum += F[b.first + m] * F[b.first + m];
        P[b.first] = sum;
      }
      // diagg (src/MOZYME/diagg.F90:16-69)
      int nij = 0;
      check_rc(mhip_mozyme_diagg1(c, &L, F.data(), P.data(), eigs.data(), eigv.data(), idiagg, &st1, fmo_dim,
                                  fmo.data(), ifmo.data(), &nij));
      st2.tiny = st1.tiny;
      st2.shift = shift;
      check_rc(mhip_mozyme_diagg2(c, &L, eigs.data(), eigv.data(), fmo.data(), ifmo.data(), nij, idiagg, &st2));
      st1.tiny = st2.tiny;
      check_rc(mhip_mozyme_density(c, &L, 0, nullptr, P.data()));
      ++idiagg;
      ++niter;
      if (three) {
        double pmax = 0.0;
        check_rc(mhip_mozyme_cnvgz(c, norbs, mp, P.data(), pold.data(), p1.data(), p2.data(), p3.data(), niter,
                                   idiag.data(), 1, &pmax));
      }
      check_rc(mhip_mozyme_fock2z(c, P.data(), F.data()));
      escf = energy(ee);
      ediff = escf - eold;
      eold = escf;
      // isitsc (src/MOZYME/isitsc.F90:16-116), single point (emin = 0)
      const double ovmax = st1.ovmax;
      if (o->heats && niter <= o->heats_cap) o->heats[niter - 1] = escf;
      if (o->print_level > 0) {
        std::fprintf(stderr, "mozyme iteration %4d  heat %16.6f  change %12.6f  ovmax %10.3e  %.3f s\n", niter,
                     escf, ediff, ovmax,
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        std::fflush(stderr);
      }
      const bool ok_now = ovmax < 5.0 * selcon && std::fabs(ediff) < selcon;
      if ((ok_now && scf1) || niter > itrmax) okscf = true;
      else scf1 = ok_now;
      if (okscf && niter > 3) break;
      // iter_for_MOZYME.F90:617-630: three-point extrapolation, then shift
      if (three) {
        if (niter % 3 == 2 && std::fabs(ediff) < 0.1) {
          three = false;
          lstart = niter;
          shift = 0.0;
        }
      } else if (ediff > 0.0 && shift < 11.0 && niter > lstart + 2) {
        shift += 2.0;
        lstart = niter;
      }
    }
    // LMOs back to the caller's arrays (same capacities)
    std::copy(occ.nc.begin(), occ.nc.end(), lmo->ncf);
    std::copy(occ.ic.begin(), occ.ic.end(), lmo->icocc);
    std::copy(occ.c.begin(), occ.c.end(), lmo->cocc);
    std::copy(vir.nc.begin(), vir.nc.end(), lmo->nce);
    std::copy(vir.ic.begin(), vir.ic.end(), lmo->icvir);
    std::copy(vir.c.begin(), vir.c.end(), lmo->cvir);
    std::copy(occ.nnc.begin(), occ.nnc.end(), const_cast<int*>(lmo->nncf));
    std::copy(occ.ncmo.begin(), occ.ncmo.end(), const_cast<int*>(lmo->ncocc));
    std::copy(vir.nnc.begin(), vir.nnc.end(), const_cast<int*>(lmo->nnce));
    std::copy(vir.ncmo.begin(), vir.ncmo.end(), const_cast<int*>(lmo->ncvir));
    if (p_out) std::copy(P.begin(), P.end(), p_out);
    if (f_out) std::copy(F.begin(), F.end(), f_out);
    res->escf = escf;
    res->ee = ee;
    res->niter = niter;
    res->converged = niter <= itrmax ? 1 : 0;
    res->ovmax = st1.ovmax;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    return 1;
  }
}

}  // extern "C"
