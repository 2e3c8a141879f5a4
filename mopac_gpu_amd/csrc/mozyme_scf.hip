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
#include "k_mozyme.h"

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

// f (device, sparse) = h + G(p) for p (device, sparse), on stream s
void mz_fock2z_on(mhip_ctx* c, hipStream_t s, const double* ps, double* fs) {
  MzFock* z = c->mz;
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

void mz_fock2z(mhip_ctx* c, const double* ps, double* fs) { mz_fock2z_on(c, c->stream, ps, fs); }

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
// The LMOs, P, F and the extrapolation arrays stay on the device for the
// whole SCF; each iteration sends the host the scalars the loop's control
// needs, tidy's per-LMO counts and diagg1's list (diagg2 orders its rotations
// on the host), and nothing else.

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
    hipStream_t s = mz_stream_of(c);
    MHIP_CHECK(hipStreamSynchronize(c->stream));  // the system upload, on the context's stream
    std::vector<int> iorbs(numat);
    for (int a = 0; a < numat; ++a) iorbs[a] = lmo->nlast[a] - lmo->nfirst[a] + 1;
    // the host side of the resident set: sizes and the LMO offsets tidy moves
    std::vector<int> nncf(lmo->nncf, lmo->nncf + lmo->nocc), ncocc(lmo->ncocc, lmo->ncocc + lmo->nocc),
        nnce(lmo->nnce, lmo->nnce + lmo->nvir), ncvir(lmo->ncvir, lmo->ncvir + lmo->nvir);
    mhip_lmo_set meta = *lmo;
    meta.nncf = nncf.data();
    meta.ncocc = ncocc.data();
    meta.nnce = nnce.data();
    meta.ncvir = ncvir.data();
    mz_resident_begin(s, lmo);
    struct EndResident {  // the LMOs go back to the caller on every exit
      hipStream_t s;
      mhip_lmo_set* L;
      bool done = false;
      ~EndResident() {
        if (!done) try { mz_resident_end(s, L); } catch (...) {}
      }
    } end_resident{s, lmo};
    // device arrays of the loop
    DevBuf P, F, pold, p1, p2, p3, idiag_d;
    P.ensure(sizeof(double) * mp);
    F.ensure(sizeof(double) * mp);
    pold.ensure(sizeof(double) * mp);
    for (DevBuf* b : {&p1, &p2, &p3}) b->ensure(sizeof(double) * norbs);
    MHIP_CHECK(hipMemsetAsync(pold.ptr, 0, sizeof(double) * mp, s));
    for (DevBuf* b : {&p1, &p2, &p3}) MHIP_CHECK(hipMemsetAsync(b->ptr, 0, sizeof(double) * norbs, s));
    struct FreeAll {
      DevBuf* bufs[7];
      ~FreeAll() {
        for (DevBuf* b : bufs) b->release();
      }
    } free_all{{&P, &F, &pold, &p1, &p2, &p3, &idiag_d}};
    // idiag: 1-based position of P(i, i) in the sparse storage
    std::vector<int> idiag(norbs);
    for (int a = 0, l = 0; a < numat; ++a) {
      int j = lmo->nijbo[a + (size_t)a * numat];
      for (int k = 1; k <= iorbs[a]; ++k) {
        j += k;
        idiag[l++] = j;
      }
    }
    idiag_d.ensure(sizeof(int) * norbs);
    MHIP_CHECK(hipMemcpyAsync(idiag_d.ptr, idiag.data(), sizeof(int) * norbs, hipMemcpyHostToDevice, s));
    int nblk = 0;
    mz_eimp_blocks(s, numat, lmo->nijbo, iorbs.data(), &nblk);
    const double* H = z->hs.as<double>();
    std::vector<double> eigs(std::max(lmo->nocc, 1), 0.0), eigv(std::max(lmo->nvir, 1), 0.0);
    const int fmo_dim = o->fmo_dim > 0 ? o->fmo_dim : (int)std::min<long long>(
        (long long)lmo->nocc * lmo->nvir, 50LL * std::max(lmo->nocc, lmo->nvir) * 64);
    std::vector<double> fmo(std::max(fmo_dim, 1));
    std::vector<int> ifmo(2 * (size_t)std::max(fmo_dim, 1)), nfmo(std::max(lmo->nvir, 1), 0);
    mhip_diagg1_state st1{};
    st1.reset = 1;
    st1.nfmo = nfmo.data();
    st1.fp32_accumulate = o->fp32_fmo ? 1 : 0;
    mhip_diagg2_state st2{};
    st2.reset = 1;
    st2.damp = o->damp > 0 ? o->damp : 1.0;
    st2.thresh = o->thresh > 0 ? o->thresh : 1.0e-13;
    st2.order_free = o->order_free ? 1 : 0;
    const double selcon = o->scfcrt > 0 ? o->scfcrt : 1.0e-2;  // scfcri (src/MOZYME/scfcri.F90)
    const int itrmax = o->itrmax > 0 ? o->itrmax : 2000;
    const double fpc9 = mhip_ev_to_kcal();
    double ee = 0.0;
    auto energy = [&]() {
      ee = mz_helecz_dev(s, numat, P.as<double>(), H, F.as<double>());
      return (ee + o->enuclr) * fpc9 + o->atheat;
    };
    // iter_for_MOZYME.F90:433-478: density from the LMOs, Fock, energy
    mz_density_dev(s, mp, lmo->nocc, 0, nullptr, P.as<double>());
    mz_fock2z_on(c, s, P.as<double>(), F.as<double>());
    double escf = energy();
    double eold = 0.0, ediff = 0.0, shift = 0.0;
    int niter = 0, idiagg = 0, lstart = 0;
    bool three = true, scf1 = false, okscf = false;
    for (;;) {
      MHIP_REQUIRE(mz_tidy_dev(s, numat, norbs, true, lmo->nocc, nncf.data(), ncocc.data(), st2.thresh) &&
                       mz_tidy_dev(s, numat, norbs, false, lmo->nvir, nnce.data(), ncvir.data(), st2.thresh),
                   "mhip_mozyme_scf: LMO storage too small for the next step (enlarge icocc/cocc/icvir/cvir)");
      MHIP_REQUIRE(mz_check_dev(s, true, lmo->nocc) <= 0.1 && mz_check_dev(s, false, lmo->nvir) <= 0.1,
                   "mhip_mozyme_scf: LMO normalization error > 0.1");
      // eimp (src/MOZYME/eimp.F90:16-45): the s-s slot of each stored
      // off-diagonal block of P holds the sum of squares of that F block
      mz_eimp_dev(s, nblk, F.as<double>(), P.as<double>());
      // diagg (src/MOZYME/diagg.F90:16-69)
      int nij = 0;
      mz_diagg1_dev(s, &meta, F.as<double>(), P.as<double>(), eigs.data(), eigv.data(), idiagg, &st1, fmo_dim,
                    fmo.data(), ifmo.data(), &nij);
      st2.tiny = st1.tiny;
      st2.shift = shift;
      mz_diagg2_dev(s, &meta, eigs.data(), eigv.data(), fmo.data(), ifmo.data(), nij, idiagg, &st2);
      st1.tiny = st2.tiny;
      mz_density_dev(s, mp, lmo->nocc, 0, nullptr, P.as<double>());
      ++idiagg;
      ++niter;
      if (three)
        mz_cnvgz_dev(s, norbs, mp, P.as<double>(), pold.as<double>(), p1.as<double>(), p2.as<double>(),
                     p3.as<double>(), niter, idiag_d.as<int>(), 1);
      mz_fock2z_on(c, s, P.as<double>(), F.as<double>());
      escf = energy();
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
    // LMOs (with the offsets tidy moved) back to the caller's arrays
    end_resident.done = true;
    mz_resident_end(s, lmo);
    if (p_out) MHIP_CHECK(hipMemcpyAsync(p_out, P.ptr, sizeof(double) * mp, hipMemcpyDeviceToHost, s));
    if (f_out) MHIP_CHECK(hipMemcpyAsync(f_out, F.ptr, sizeof(double) * mp, hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipStreamSynchronize(s));
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
