// Context, system upload, the device-resident SCF driver and the C ABI.
//
// The driver restates the control flow of iter (src/SCF/iter.F90:16-1107) for
// a closed-shell RHF molecule: level shift, Fock build, energy, convergence
// test, Pulay DIIS, full diagonalization, density, cnvg extrapolation.  Every
// matrix stays in HBM; per iteration only a handful of scalars (energy, PL,
// two frontier eigenvalues) come back to the host, which is where the
// reference keeps its control logic too.
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"
#include "k_eigen.h"
#include "k_fock.h"
#include "k_guess.h"
#include "k_occvirt.h"
#include "k_pseudo.h"

#include "ctx.h"


namespace mhip {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const std::string& get_error() { return g_err; }

namespace {

template <class F>
int guarded(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    return 1;
  } catch (const std::exception& e) {
    set_error(e.what());
    return 1;
  }
}


__global__ void k_diag_add_index(int n, double* p, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[(long long)i * (i + 1) / 2 + i] += (double)(i + 1) * scale;
}

__global__ void k_set_diag_packed(int n, double* p, const double* d, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[(long long)i * (i + 1) / 2 + i] = d[i] * scale;
}

__global__ void k_get_diag_packed(int n, const double* p, double* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  d[i] = p[(long long)i * (i + 1) / 2 + i];
}

// out[0] = max eig over the occupied columns, out[1] = min over the virtual
// ones (the frontier levels whatever the column order)
__global__ __launch_bounds__(256) void k_frontier(int n, int nocc, const double* __restrict__ eig,
                                                  double* __restrict__ out) {
  __shared__ double hi[256], lo[256];
  double h = -1.0e300, l = 1.0e300;
  for (int i = threadIdx.x; i < n; i += 256) {
    if (i < nocc) h = fmax(h, eig[i]);
    else l = fmin(l, eig[i]);
  }
  hi[threadIdx.x] = h;
  lo[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + w]);
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = hi[0];
    out[1] = lo[0];
  }
}

// identity on the padding columns/rows [n, npad), zero elsewhere
__global__ void k_pad_identity(int n, int npad, double* __restrict__ c) {
  const long long len = (long long)npad * npad;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e % npad), col = (int)(e / npad);
    c[e] = (r == col && r >= n) ? 1.0 : 0.0;
  }
}

}  // namespace

void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
  b.ensure(bytes ? bytes : 8);
  if (bytes) MHIP_CHECK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s));
}

void copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  MHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

// Build the pair tables for the Fock kernels from nfirst/nlast (1-based).
void build_system(mhip_ctx* c, int numat, const int* nfirst, const int* nlast, int norbs,
                         long long n2elec, const double* h, const double* w) {
  hipStream_t s = c->stream;
  c->numat = numat;
  c->norbs = norbs;
  c->mpack = mpack_of(norbs);
  c->n2elec = n2elec;
  std::vector<int> nf0(numat), no(numat), heavy, light, nhb(numat), nlb(numat);
  std::vector<long long> woff(numat), wcum(numat), w1off(numat), ccol(numat);
  long long kr = 0, wc = 0, cc = 0;
  for (int i = 0; i < numat; ++i) {
    nf0[i] = nfirst[i] - 1;
    no[i] = nlast[i] - nfirst[i] + 1;
    MHIP_REQUIRE(no[i] == 1 || no[i] == 4, "only s (H) and sp (C, N, O) atoms are supported");
    MHIP_REQUIRE(i == 0 || nf0[i] == nf0[i - 1] + no[i - 1], "orbitals must be contiguous by atom");
  }
  MHIP_REQUIRE(nf0[numat - 1] + no[numat - 1] == norbs, "nlast(numat) must equal norbs");
  c->nf0_h = nf0;
  c->no_h = no;
  for (int i = 0; i < numat; ++i) {
    const int lim = no[i] * (no[i] + 1) / 2;
    woff[i] = kr;
    wcum[i] = wc;
    kr += (long long)lim * wc;   // pairs (i, j<i)
    w1off[i] = kr;
    kr += (long long)lim * lim;  // one-centre block
    wc += lim;
    nhb[i] = (int)heavy.size();
    nlb[i] = (int)light.size();
    (no[i] == 4 ? heavy : light).push_back(i);
    ccol[i] = cc;
    cc += (long long)(numat - 1 - i) * lim;
  }
  MHIP_REQUIRE(kr == n2elec, "n2elec does not match the hcore integral layout");
  upload(c->w, w, sizeof(double) * n2elec, s);
  upload(c->h, h, sizeof(double) * c->mpack, s);
  upload(c->nf0, nf0.data(), sizeof(int) * numat, s);
  upload(c->norb, no.data(), sizeof(int) * numat, s);
  upload(c->woff, woff.data(), sizeof(long long) * numat, s);
  upload(c->wcum, wcum.data(), sizeof(long long) * numat, s);
  upload(c->w1off, w1off.data(), sizeof(long long) * numat, s);
  upload(c->ccol, ccol.data(), sizeof(long long) * numat, s);
  upload(c->heavy, heavy.data(), sizeof(int) * heavy.size(), s);
  upload(c->light, light.data(), sizeof(int) * light.size(), s);
  upload(c->nhb, nhb.data(), sizeof(int) * numat, s);
  upload(c->nlb, nlb.data(), sizeof(int) * numat, s);
  c->jrow.ensure(sizeof(double) * 10 * numat);
  c->jcol.ensure(sizeof(double) * (cc > 0 ? cc : 1));
  SysDev& d = c->sys;
  d.numat = numat;
  d.norbs = norbs;
  d.n2elec = n2elec;
  d.nf0 = c->nf0.as<int>();
  d.norb = c->norb.as<int>();
  d.woff = c->woff.as<long long>();
  d.wcum = c->wcum.as<long long>();
  d.w1off = c->w1off.as<long long>();
  d.ccol = c->ccol.as<long long>();
  d.heavy = c->heavy.as<int>();
  d.light = c->light.as<int>();
  d.nheavy_before = c->nhb.as<int>();
  d.nlight_before = c->nlb.as<int>();
  d.w = c->w.as<double>();
  // SCF work buffers
  const long long mp = c->mpack;
  c->plan.build(norbs);
  c->npad = c->plan.npad;
  const size_t dense = sizeof(double) * (size_t)c->npad * c->npad;
  for (DevBuf* b : {&c->p, &c->pa, &c->f, &c->fcopy, &c->pold}) b->ensure(sizeof(double) * mp);
  c->pold2.ensure(sizeof(double) * norbs);
  for (DevBuf* b : {&c->fd, &c->A, &c->C, &c->C2, &c->T}) b->ensure(dense);
  c->eig.ensure(sizeof(double) * c->npad);
  c->scal.ensure(sizeof(double) * 64);
  c->partial.ensure(sizeof(double) * std::max<size_t>(norbs, dots_partial_doubles(mp, 11)) + 64);
  c->have_c = false;
  c->have_system = true;
  c->have_guess = false;
  MHIP_CHECK(hipStreamSynchronize(s));
}

FockArgs fock_args(mhip_ctx* c, int mode, double shift) {
  FockArgs a;
  a.sys = c->sys;
  a.h = c->h.as<double>();
  a.p = c->p.as<double>();
  a.pa = c->pa.as<double>();
  a.fin = c->f.as<double>();
  a.f = c->f.as<double>();
  a.jrow = c->jrow.as<double>();
  a.jcol = c->jcol.as<double>();
  a.shift = shift;
  a.init_mode = mode;
  return a;
}

// Device -> caller's (pageable) host memory: DMA into the context's pinned
// staging buffer, then the host copy split over a few threads (a single
// thread's copy of a 38 MB density took longer than the DMA).
static void download(mhip_ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  if (c->hstage_bytes < bytes) {
    if (c->hstage) MHIP_CHECK(hipHostFree(c->hstage));
    c->hstage = nullptr;
    c->hstage_bytes = 0;
    MHIP_CHECK(hipHostMalloc(&c->hstage, bytes));
    c->hstage_bytes = bytes;
  }
  MHIP_CHECK(hipMemcpyAsync(c->hstage, src, bytes, hipMemcpyDeviceToHost, c->stream));
  MHIP_CHECK(hipStreamSynchronize(c->stream));
  const char* from = static_cast<const char*>(c->hstage);
  char* to = static_cast<char*>(dst);
  const size_t chunk = size_t(4) << 20;
  const int nthr = (int)std::min<size_t>(8, (bytes + chunk - 1) / chunk);
  if (nthr <= 1) {
    std::memcpy(to, from, bytes);
    return;
  }
  const size_t part = (bytes + nthr - 1) / nthr;
  std::vector<std::thread> pool;
  for (int t = 1; t < nthr; ++t) {
    const size_t b0 = t * part, b1 = std::min(bytes, b0 + part);
    if (b0 < b1) pool.emplace_back([=] { std::memcpy(to + b0, from + b0, b1 - b0); });
  }
  std::memcpy(to, from, std::min(bytes, part));
  for (auto& th : pool) th.join();
}

// eigenvectors_LAPACK on the device copy of a packed matrix: adds i*1e-10 to
// the diagonal in place, then solves; ctx C/eig hold the result (sorted).
static void eigen_packed(mhip_ctx* c, double* fpacked, bool perturb, double tol = 1.0e-12,
                         JacobiPlan* plan_in = nullptr) {
  JacobiPlan& plan = plan_in ? *plan_in : c->plan;
  hipStream_t s = c->stream;
  const int n = c->norbs, np = c->npad;
  const double pad_base = 1.0e4;
  if (perturb) {
    hipLaunchKernelGGL(k_diag_add_index, dim3((n + 255) / 256), dim3(256), 0, s, n, fpacked, 1.0e-10);
    MHIP_LAUNCH_CHECK();
  }
  double* A = c->A.as<double>();
  double* C = c->C.as<double>();
  if (c->have_c) {
    unpack_sym(s, n, fpacked, c->fd.as<double>(), np, np, pad_base);
    gemm(s, 'N', 'N', np, np, np, 1.0, c->fd.as<double>(), np, C, np, 0.0, c->T.as<double>(), np);
    gemm(s, 'T', 'N', np, np, np, 1.0, C, np, c->T.as<double>(), np, 0.0, A, np);
  } else {
    unpack_sym(s, n, fpacked, A, np, np, pad_base);
    jacobi_set_identity(s, np, C, np);
  }
  int sweeps = 0;
  const int rc = jacobi_sweeps(s, plan, A, C, tol, 40, c->hscal + 32,
                               c->scal.as<double>() + 32, &sweeps);
  MHIP_REQUIRE(rc == 0, "block-Jacobi eigensolver did not converge in 40 sweeps");
  c->last_sweeps = sweeps;
  c->total_sweeps += sweeps;
  jacobi_sort(s, plan, A, C, c->C2.as<double>(), c->eig.as<double>());
  std::swap(c->C, c->C2);
  c->have_c = true;
}

// ---- Pulay DIIS (src/SCF/pulay.F90:16-166) -------------------------------------
struct PulayState {
  int mfock = 0, nfock = 0, lfock = 0;
  bool start = true;
  double emat[11 * 11];
};

// Full-pivot Gauss-Jordan inverse with determinant (the algorithm of
// src/matrix/osinv.F90); a is n x n column-major.
static void osinv(double* a, int n, double* d_out) {
  const double tol = 1.0e-8;
  int l[16], m[16];
  double d = 1.0;
  for (int k = 0; k < n; ++k) {
    l[k] = k; m[k] = k;
    double biga = a[k * n + k];
    for (int j = k; j < n; ++j)
      for (int i = k; i < n; ++i)
        if (std::fabs(biga) < std::fabs(a[j * n + i])) { biga = a[j * n + i]; l[k] = i; m[k] = j; }
    int j = l[k];
    if (j > k)
      for (int i = 0; i < n; ++i) {
        const double holo = -a[i * n + k];
        a[i * n + k] = a[i * n + j];
        a[i * n + j] = holo;
      }
    int i = m[k];
    if (i > k)
      for (int jj = 0; jj < n; ++jj) {
        const double holo = -a[k * n + jj];
        a[k * n + jj] = a[i * n + jj];
        a[i * n + jj] = holo;
      }
    if (std::fabs(biga) < tol) { *d_out = 0.0; return; }
    for (int ii = 0; ii < n; ++ii)
      if (ii != k) a[k * n + ii] /= -biga;
    for (int ii = 0; ii < n; ++ii) {
      if (ii == k) continue;
      const double holo = a[k * n + ii];
      for (int jj = 0; jj < n; ++jj)
        if (jj != k) a[jj * n + ii] += holo * a[jj * n + k];
    }
    for (int jj = 0; jj < n; ++jj)
      if (jj != k) a[jj * n + k] /= biga;
    d = std::min(d * biga, 1.0e10);
    a[k * n + k] = 1.0 / biga;
  }
  for (int k = n - 2; k >= 0; --k) {
    int i = l[k];
    if (i > k)
      for (int j = 0; j < n; ++j) {
        const double holo = a[k * n + j];
        a[k * n + j] = -a[i * n + j];
        a[i * n + j] = holo;
      }
    int j = m[k];
    if (j > k)
      for (int ii = 0; ii < n; ++ii) {
        const double holo = a[ii * n + k];
        a[ii * n + k] = -a[ii * n + j];
        a[ii * n + j] = holo;
      }
  }
  *d_out = d;
}

static void pulay_step(mhip_ctx* c, PulayState& st, double* pl_out) {
  hipStream_t s = c->stream;
  const int n = c->norbs, np = c->npad;
  const long long mp = c->mpack;
  const int npulay = 10;
  if (!c->fhist.ptr || c->fhist.bytes < sizeof(double) * npulay * mp) {
    c->fhist.ensure(sizeof(double) * npulay * mp);
    c->ehist.ensure(sizeof(double) * npulay * mp);
    c->coef.ensure(sizeof(double) * 16);
  }
  if (st.start) {
    st.mfock = npulay;
    st.nfock = 1;
    st.lfock = 1;
    st.start = false;
  } else {
    if (st.nfock < st.mfock) st.nfock += 1;
    st.lfock = (st.lfock != st.mfock) ? st.lfock + 1 : 1;
  }
  const int nfock = st.nfock, lfock = st.lfock;
  double* fh = c->fhist.as<double>();
  double* eh = c->ehist.as<double>();
  copy_d2d(fh + (size_t)(lfock - 1) * mp, c->f.ptr, sizeof(double) * mp, s);
  // FP - PF (P = alpha density), packed as the reference packs it
  unpack_sym(s, n, c->pa.as<double>(), c->fd.as<double>(), np, np, 0.0);
  unpack_sym(s, n, c->f.as<double>(), c->T.as<double>(), np, np, 0.0);
  gemm(s, 'N', 'N', n, n, n, 1.0, c->fd.as<double>(), np, c->T.as<double>(), np, 0.0,
       c->A.as<double>(), np);
  pack_antisym(s, n, c->A.as<double>(), np, eh + (size_t)(lfock - 1) * mp);
  double* dev_dots = c->scal.as<double>();
  dots(s, mp, nfock, eh, mp, eh + (size_t)(lfock - 1) * mp, dev_dots, c->partial.as<double>());
  MHIP_CHECK(hipMemcpyAsync(c->hscal, dev_dots, sizeof(double) * nfock, hipMemcpyDeviceToHost, s));
  MHIP_CHECK(hipStreamSynchronize(s));
  const int ld = npulay + 1;
  auto EM = [&](int i, int j) -> double& { return st.emat[(j - 1) * ld + (i - 1)]; };
  const int nf1 = nfock + 1;
  for (int i = 1; i <= nfock; ++i) {
    EM(nf1, i) = -1.0;
    EM(i, nf1) = -1.0;
    EM(lfock, i) = c->hscal[i - 1];
    EM(i, lfock) = c->hscal[i - 1];
  }
  *pl_out = EM(lfock, lfock) / (double)mp;
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
  double d;
  osinv(evec, nf1, &d);
  if (std::fabs(d) < 1.0e-6) {
    st.start = true;
    return;
  }
  if (nfock < 2) return;
  const int il = nfock * nf1;
  double coeffs[16];
  for (int i = 0; i < nfock; ++i) coeffs[i] = -evec[il + i];
  MHIP_CHECK(hipMemcpyAsync(c->coef.ptr, coeffs, sizeof(double) * nfock, hipMemcpyHostToDevice, s));
  lincomb(s, mp, nfock, c->coef.as<double>(), fh, mp, c->f.as<double>());
}

static double read_scalar(mhip_ctx* c, const double* dev) {
  MHIP_CHECK(hipMemcpyAsync(c->hscal, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MHIP_CHECK(hipStreamSynchronize(c->stream));
  return c->hscal[0];
}

static double energy(mhip_ctx* c) {
  helect(c->stream, c->norbs, c->pa.as<double>(), c->h.as<double>(), c->f.as<double>(), 2.0,
         c->partial.as<double>(), c->scal.as<double>() + 16);
  return read_scalar(c, c->scal.as<double>() + 16);
}

static void density(mhip_ctx* c, int nocc, bool split) {
  const int n = c->norbs;
  if (split) {
    c->split_scratch.ensure(syrk_split_scratch_bytes(n, nocc));
    syrk_packed_split_bf16(c->stream, n, nocc, 2.0, c->C.as<double>(), c->npad, c->p.as<double>(),
                           c->split_scratch.ptr);
  } else {
    syrk_packed(c->stream, n, nocc, 2.0, c->C.as<double>(), c->npad, c->p.as<double>());
  }
}

// Host tables of the fragment start; returns the number of fragments, or 0
// when the map is absent or unusable (then the first iteration diagonalizes F
// in full, as the reference does).
static int setup_fragments(mhip_ctx* c, const mhip_scf_opts* o) {
  if (!o->fragment || o->pseudo_diag != 2 || o->block_start < 2 || !o->pdiag) return 0;
  const int numat = c->numat;
  int nfrag = 0;
  for (int a = 0; a < numat; ++a) {
    if (o->fragment[a] < 0) return 0;
    nfrag = std::max(nfrag, o->fragment[a] + 1);
  }
  std::vector<std::vector<int>> orbs(nfrag);
  for (int a = 0; a < numat; ++a)
    for (int k = 0; k < c->no_h[a]; ++k) orbs[o->fragment[a]].push_back(c->nf0_h[a] + k);
  std::vector<int> ptr(nfrag + 1, 0), orb, focc(nfrag), ooff(nfrag), voff(nfrag);
  int nocc_sum = 0, nvirt_sum = 0;
  for (int f = 0; f < nfrag; ++f) {
    const int nf = (int)orbs[f].size();
    if (nf == 0 || nf > kMaxFragOrbs) return 0;
    double el = 0.0;
    for (int i : orbs[f]) el += o->pdiag[i];
    const long long ne = std::llround(el);
    if (std::fabs(el - (double)ne) > 1e-6 || (ne & 1) || ne / 2 > nf) return 0;
    focc[f] = (int)(ne / 2);
    ptr[f + 1] = ptr[f] + nf;
    orb.insert(orb.end(), orbs[f].begin(), orbs[f].end());
    ooff[f] = nocc_sum;
    nocc_sum += focc[f];
  }
  if (nocc_sum != o->nclose) return 0;
  for (int f = 0; f < nfrag; ++f) {
    voff[f] = o->nclose + nvirt_sum;
    nvirt_sum += (int)orbs[f].size() - focc[f];
  }
  hipStream_t s = c->stream;
  upload(c->frag_ptr, ptr.data(), sizeof(int) * ptr.size(), s);
  upload(c->frag_orb, orb.data(), sizeof(int) * orb.size(), s);
  upload(c->frag_occ, focc.data(), sizeof(int) * nfrag, s);
  upload(c->frag_ooff, ooff.data(), sizeof(int) * nfrag, s);
  upload(c->frag_voff, voff.data(), sizeof(int) * nfrag, s);
  MHIP_CHECK(hipStreamSynchronize(s));  // the host vectors go out of scope
  return nfrag;
}

void run_scf(mhip_ctx* c, const mhip_scf_opts* o, mhip_scf_result* r, double* p_out,
             double* f_out, double* c_out, double* eig_out) {
  MHIP_REQUIRE(c->have_system, "mhip_upload_system must be called before mhip_scf");
  hipStream_t s = c->stream;
  const int n = c->norbs;
  const long long mp = c->mpack;
  const double enrgy = 23.060547830619029;
  const auto t0 = std::chrono::steady_clock::now();
  double* P = c->p.as<double>();
  double* Pa = c->pa.as<double>();
  double* pold = c->pold.as<double>();
  c->have_c = false;
  c->total_sweeps = 0;
  // ---- initialization (iter.F90:91-251)
  double ten = 10.0, bshift = -80.0, shift = 1.0, shfto = 0.0, shftbo = 0.0;
  double shiftb = 0.0, shfmax = 20.0, plchek = 0.005, pl = 1.0, trans = 0.2;
  const int itrmax = o->itrmax > 0 ? o->itrmax : 2000;
  const bool okpuly = o->pulay != 0;
  bool oknewd = std::fabs(bshift) < 0.001;
  bool newdg = false;
  if (std::fabs(bshift) > 1e-20) ten = bshift;
  const int nclose = o->nclose;
  const double w1 = nclose / (nclose + 1.0e-6 + nclose);
  if (o->start_from_guess) {
    MHIP_REQUIRE(c->have_guess, "start_from_guess needs mhip_guess_water on the uploaded system");
    copy_d2d(P, c->pguess.ptr, sizeof(double) * mp, s);
    axpby(s, mp, 0.5, P, 0.0, Pa);
    copy_d2d(pold, P, sizeof(double) * mp, s);
  } else if (o->p_init) {
    MHIP_CHECK(hipMemcpyAsync(P, o->p_init, sizeof(double) * mp, hipMemcpyHostToDevice, s));
    axpby(s, mp, 0.5, P, 0.0, Pa);
    copy_d2d(pold, P, sizeof(double) * mp, s);
  } else {
    MHIP_REQUIRE(o->pdiag != nullptr, "pdiag or p_init is required");
    fill(s, P, mp, 0.0);
    fill(s, Pa, mp, 0.0);
    MHIP_CHECK(hipMemcpyAsync(c->pold2.ptr, o->pdiag, sizeof(double) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_set_diag_packed, dim3((n + 255) / 256), dim3(256), 0, s, n, P,
                       c->pold2.as<double>(), 1.0);
    hipLaunchKernelGGL(k_set_diag_packed, dim3((n + 255) / 256), dim3(256), 0, s, n, Pa,
                       c->pold2.as<double>(), w1 < 1e-6 ? 0.5 : w1);
    MHIP_LAUNCH_CHECK();
    copy_d2d(pold, Pa, sizeof(double) * mp, s);
  }
  hipLaunchKernelGGL(k_get_diag_packed, dim3((n + 255) / 256), dim3(256), 0, s, n, pold,
                     c->pold2.as<double>());
  MHIP_LAUNCH_CHECK();
  double scfcrt = o->scfcrt > 0 ? o->scfcrt : 1.0e-4;
  scfcrt = std::max(scfcrt, 1e-12);
  const double selcon = std::min(std::min(scfcrt * 100.0, 0.1), scfcrt);
  const double pltest = 0.05 * std::sqrt(std::fabs(selcon));
  int iredy = 1, niter = 0;
  bool ready = false, converged = false;
  double eold = 100.0, diff = 0.0, sellim = 0.0, escf = 0.0, ee = 0.0;
  const int ihomo = std::max(nclose, 1);
  double eig_homo = 0.0, eig_lumo = 0.0;
  PulayState pst;
  int nfull = 0, npseudo = 0, block_steps_since_diag = 0, nfragdiag = 0, n_split = 0, fp64_block_steps = 0;
  const int nfrag = setup_fragments(c, o);
  c->frontier.ensure(sizeof(double) * 2);
  for (;;) {
    niter += 1;
    int mode = kInitH;
    if (std::fabs(shift) > 1e-10 && bshift != 0.0) {
      if (niter > 1) {
        if (newdg) {
          if (diff > 0.0) {
            shift = 1.0;
            if (diff > 1) newdg = false;
          } else {
            shift = -0.1;
          }
        } else {
          shift = (ihomo < n) ? ten + eig_lumo - eig_homo + shift : 0.0;
        }
        if (diff > 0.0) {
          if (shift > 4.0) shfmax = 4.5;
          if (shift > shfmax) shfmax = std::max(shfmax - 0.5, 0.0);
        }
        shift = std::max(-20.0, std::min(shfmax, shift));
        if (std::fabs(shift - shfmax) < 1e-5) shfmax += 0.01;
        if (okpuly || std::fabs(bshift - 4.44) < 1e-5) {
          shift = -8.0;
          if (newdg) shift = 0.0;
        }
      }
      if (pl > plchek) {
        shftbo = shiftb;
        shfto = shift;
      } else {
        shiftb = shftbo;
        shift = shfto;
      }
      mode = kInitShift;
    } else if (o->last == 0 && niter < 2) {
      mode = kInitRandom;
    }
    bool recompute = false;
    for (;;) {  // label 320
      fock_build(s, fock_args(c, recompute ? kInitH : mode, shift));
      if (niter >= itrmax) break;
      ee = energy(c);
      MHIP_REQUIRE(std::isfinite(ee), "SCF produced a non-finite electronic energy");
      escf = (ee + o->enuclr) * enrgy + o->atheat;
      diff = escf - eold;
      if (diff > 0) ten = ten - 1.0;
      else ten = ten * 0.975 + 0.05;
      sellim = std::max(selcon, 1e-15 * std::max(std::fabs(ee), 1.0));
      if (niter > 4 && (pl == 0.0 || (pl < pltest && std::fabs(diff) < sellim)) && ready) {
        if (std::fabs(shift) < 1e-10) {
          converged = true;
          break;
        }
        shift = 0.0;
        shiftb = 0.0;
        recompute = true;
        continue;
      }
      break;
    }
    if (converged || niter >= itrmax) break;
    ready = iredy > 0 && (std::fabs(diff) < sellim * 10.0 || pl == 0.0);
    iredy += 1;
    eold = escf;
    if (newdg && okpuly && iredy > 1) pulay_step(c, pst, &pl);
    // precision form: split-BF16 contractions while far from self-consistency,
    // FP64 for the iterations that decide convergence
    const bool split_now = o->precision == 1 && pl > 100.0 * pltest;
    n_split += split_now ? 1 : 0;
    // block steps before newdg only from good first eigenvectors: the
    // fragment blocks' or those of a supplied starting density.  From the
    // standard diagonal guess the first diagonalization is far from the
    // answer and first-order occupied-virtual steps from it need not settle,
    // so those wait for newdg (PL < TRANS), as iter's pseudo-diagonalization
    // does (iter.F90:1039).
    const bool early_blocks = o->pseudo_diag == 2 && o->block_start >= 1 &&
                              (nfragdiag > 0 || o->start_from_guess || o->p_init != nullptr);
    const bool use_pseudo = o->pseudo_diag && (newdg || early_blocks) && c->have_c && !okpuly;
    if (!c->have_c && nfrag > 0) {
      // fragment start: block-diagonal eigenvectors of F's fragment blocks
      c->ov.ensure(n, nclose, c->npad);
      unpack_sym(s, n, c->f.as<double>(), c->ov.fd.as<double>(), c->npad, n, 0.0);
      hipLaunchKernelGGL(k_pad_identity, dim3(2048), dim3(256), 0, s, n, c->npad, c->C.as<double>());
      MHIP_LAUNCH_CHECK();
      fragment_eigen(s, nfrag, c->ov.fd.as<double>(), c->npad, c->frag_ptr.as<int>(),
                     c->frag_orb.as<int>(), c->frag_occ.as<int>(), c->frag_ooff.as<int>(),
                     c->frag_voff.as<int>(), c->C.as<double>(), c->eig.as<double>());
      c->have_c = true;
      block_steps_since_diag += 1;
      ++nfragdiag;
    } else if (use_pseudo) {
      if (o->pseudo_diag == 2) {
        // Bjorck error after k steps ~ |B^T B|^(2^k), and each iteration's
        // step also squares what the last one left: a second step only
        // while PL is large (the rotations are big)
        const int nbj = pl > 0.05 ? 2 : 1;
        // The FP64 iterations, which decide convergence, also take out the
        // virtual columns' overlap with the occupied ones: the first of them
        // (the overlap the large early rotations left) and every fourth after
        // it (each later step adds only |E||B|, the blocks' Bjorck residue
        // times the rotation size)
        const bool project = !split_now && (fp64_block_steps++ % 4 == 0);
        occvirt_step(s, c->ov, n, c->npad, c->f.as<double>(), c->C.as<double>(), nclose,
                     c->eig.as<double>(), split_now, nbj, project);
        block_steps_since_diag += 1;
      } else {
        pseudo_diag_device(s, c->pseudo, n, c->npad, c->f.as<double>(), c->C.as<double>(), nclose,
                           c->eig.as<double>(), nullptr);
      }
      ++npseudo;
    } else {
      if (block_steps_since_diag > 0) {
        // the warm start needs orthonormal columns to working precision
        bjorck_step(s, c->ov, n, c->npad, c->C.as<double>(), 0, nclose);
        orthogonalize_virtual(s, c->ov, n, c->npad, c->C.as<double>(), nclose, 1);
        block_steps_since_diag = 0;
      }
      double tol = c->tol_iter;
      if (o->adaptive_eigen && pl > 10.0 * pltest) {
        // far from self-consistency the eigenvectors only need to be as good
        // as the density they will produce
        tol = std::max(c->tol_iter, std::min(1.0e-5, 1.0e-3 * pl));
      }
      eigen_packed(c, c->f.as<double>(), true, tol);
      ++nfull;
    }
    density(c, nclose, split_now);
    if (!(newdg && okpuly)) {
      cnvg(s, n, P, pold, c->pold2.as<double>(), niter, c->scal.as<double>());
      // pl, plus the two frontier eigenvalues for the next level shift
      MHIP_CHECK(hipMemcpyAsync(c->hscal, c->scal.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (ihomo < n) {
      hipLaunchKernelGGL(k_frontier, dim3(1), dim3(256), 0, s, n, nclose, c->eig.as<double>(),
                         c->frontier.as<double>());
      MHIP_LAUNCH_CHECK();
      MHIP_CHECK(hipMemcpyAsync(c->hscal + 1, c->frontier.ptr, 2 * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    }
    MHIP_CHECK(hipStreamSynchronize(s));
    if (!(newdg && okpuly)) pl = c->hscal[0];
    if (ihomo < n) {
      eig_homo = c->hscal[1];
      eig_lumo = c->hscal[2];
    }
    axpby(s, mp, 0.5, P, 0.0, Pa);
    oknewd = pl < sellim || oknewd;
    newdg = (pl < trans && oknewd) || newdg;
    if (pl < trans * 0.3333) oknewd = true;
  }
  // ---- exit (iter.F90:1052-1106)
  ee = energy(c);
  JacobiPlan* exit_plan = nullptr;
  if (block_steps_since_diag > 0) {
    // orthonormal to FP64 precision before the exit diagonalization: two
    // Bjorck steps on the occupied block, the virtual block projected off
    // the occupied one, two Bjorck steps on it
    bjorck_step(s, c->ov, n, c->npad, c->C.as<double>(), 0, nclose);
    bjorck_step(s, c->ov, n, c->npad, c->C.as<double>(), 0, nclose);
    orthogonalize_virtual(s, c->ov, n, c->npad, c->C.as<double>(), nclose, 2);
    // C^T F C is block diagonal between the occupied and virtual columns up
    // to the converged FMO: diagonalize the two blocks as independent
    // problems in the same steps
    if (c->plan_split.n != n || c->plan_split.split != nclose) c->plan_split.build(n, nclose);
    c->plan_split.max_inner = c->plan.max_inner;
    c->plan_split.timing = c->plan.timing;
    if (c->plan_split.grouped && o->pseudo_diag == 2) exit_plan = &c->plan_split;
  }
  copy_d2d(c->fcopy.ptr, c->f.ptr, sizeof(double) * mp, s);
  eigen_packed(c, c->fcopy.as<double>(), true, 1.0e-12, exit_plan);
  ++nfull;
  if (o->last == 1) phase_lock(s, n, c->C.as<double>(), c->npad);
  MHIP_CHECK(hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  escf = (ee + o->enuclr) * enrgy + o->atheat;
  if (r) {
    r->escf = escf;
    r->ee = ee;
    r->niter = niter;
    r->converged = converged ? 1 : 0;
    r->n_full_diag = nfull;
    r->n_pseudo_diag = npseudo;
    r->n_fragment_diag = nfragdiag;
    r->n_split_iter = n_split;
    r->seconds = std::chrono::duration<double>(t1 - t0).count();
    r->jacobi_sweeps = c->total_sweeps;
  }
  if (p_out) download(c, p_out, P, sizeof(double) * mp);
  if (f_out) download(c, f_out, c->f.ptr, sizeof(double) * mp);
  if (eig_out) download(c, eig_out, c->eig.ptr, sizeof(double) * n);
  if (c_out)
    MHIP_CHECK(hipMemcpy2D(c_out, sizeof(double) * n, c->C.ptr, sizeof(double) * c->npad,
                           sizeof(double) * n, n, hipMemcpyDeviceToHost));
}

}  // namespace mhip

// =============================================================================
// C ABI
// =============================================================================
using namespace mhip;

extern "C" {

const char* mhip_last_error(void) { return get_error().c_str(); }

int mhip_init(int device, mhip_ctx** out) {
  return guarded([&] {
    MHIP_REQUIRE(out != nullptr, "mhip_init: null output pointer");
    int ndev = 0;
    MHIP_CHECK(hipGetDeviceCount(&ndev));
    MHIP_REQUIRE(device >= 0 && device < ndev, "mhip_init: no such HIP device");
    MHIP_CHECK(hipSetDevice(device));
    auto* c = new mhip_ctx();
    c->device = device;
    MHIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    MHIP_CHECK(hipHostMalloc((void**)&c->hscal, sizeof(double) * 128, hipHostMallocDefault));
    *out = c;
  });
}

int mhip_finalize(mhip_ctx* c) {
  if (!c) return 0;
  return guarded([&] {
    MHIP_CHECK(hipSetDevice(c->device));
    MHIP_CHECK(hipStreamSynchronize(c->stream));
    for (DevBuf* b : {&c->w, &c->h, &c->nf0, &c->norb, &c->woff, &c->wcum, &c->w1off, &c->ccol,
                      &c->heavy, &c->light, &c->nhb, &c->nlb, &c->jrow, &c->jcol, &c->p, &c->pa,
                      &c->f, &c->fcopy, &c->pold, &c->pold2, &c->fd, &c->A, &c->C, &c->C2, &c->T,
                      &c->eig, &c->scal, &c->partial, &c->fhist, &c->ehist, &c->coef,
                      &c->split_scratch, &c->pguess})
      b->release();
    for (DevBuf* b : {&c->gw.xyz, &c->gw.nat, &c->gw.mol, &c->gw.side, &c->gw.waters, &c->gw.R,
                      &c->gw.geo, &c->gw.atomf, &c->gw.pa, &c->gw.pb, &c->gw.typ, &c->gw.nf0,
                      &c->gw.no, &c->gw.norm})
      b->release();
    for (DevBuf* b : {&c->frag_ptr, &c->frag_orb, &c->frag_occ, &c->frag_ooff, &c->frag_voff,
                      &c->frontier})
      b->release();
    c->plan.release();
    c->plan_split.release();
    c->pseudo.release();
    c->ov.release();
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->hstage) (void)hipHostFree(c->hstage);
    mhip::release_split_scratch(c->stream);
    mhip::release_splitk_scratch(c->stream);
    mhip::mz_release(c->mz);
    c->mz = nullptr;
    (void)hipStreamDestroy(c->stream);
    delete c;
  });
}

int mhip_device_info(mhip_ctx* c, char* name, int name_len, double* hbm_gib, int* n_cu) {
  return guarded([&] {
    hipDeviceProp_t prop;
    MHIP_CHECK(hipGetDeviceProperties(&prop, c ? c->device : 0));
    if (name && name_len > 0) {
      std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (hbm_gib) *hbm_gib = (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
    if (n_cu) *n_cu = prop.multiProcessorCount;
  });
}

int mhip_upload_system(mhip_ctx* c, int numat, const int* nfirst, const int* nlast, int norbs,
                       long long n2elec, const double* h, const double* w) {
  return guarded([&] {
    MHIP_REQUIRE(c && nfirst && nlast && h && w && numat > 0 && norbs > 0, "mhip_upload_system: bad arguments");
    MHIP_CHECK(hipSetDevice(c->device));
    build_system(c, numat, nfirst, nlast, norbs, n2elec, h, w);
  });
}

int mhip_density(mhip_ctx* c, const double* cvec, int ndubl, double occ, double* p) {
  return guarded([&] {
    MHIP_REQUIRE(c && c->have_system, "mhip_density: no system uploaded");
    MHIP_REQUIRE(ndubl >= 0 && ndubl <= c->norbs, "mhip_density: ndubl out of range");
    const int n = c->norbs;
    if (cvec) {
      MHIP_CHECK(hipMemcpy2DAsync(c->C.ptr, sizeof(double) * c->npad, cvec, sizeof(double) * n,
                                  sizeof(double) * n, n, hipMemcpyHostToDevice, c->stream));
      c->have_c = false;
    }
    syrk_packed(c->stream, n, ndubl, occ, c->C.as<double>(), c->npad, c->p.as<double>());
    if (p) MHIP_CHECK(hipMemcpyAsync(p, c->p.ptr, sizeof(double) * c->mpack, hipMemcpyDeviceToHost, c->stream));
    MHIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int mhip_fock2(mhip_ctx* c, double* f, const double* ptot, const double* p, int mode) {
  return guarded([&] {
    MHIP_REQUIRE(c && c->have_system && f && ptot && p, "mhip_fock2: bad arguments");
    MHIP_REQUIRE(mode == 1 || mode == 2, "mhip_fock2: mode must be 1 or 2");
    MHIP_REQUIRE(mode == 2, "mhip_fock2: mode 1 (no one-centre terms) is not used on the RHF path");
    const size_t b = sizeof(double) * c->mpack;
    hipStream_t s = c->stream;
    MHIP_CHECK(hipMemcpyAsync(c->f.ptr, f, b, hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(c->p.ptr, ptot, b, hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(c->pa.ptr, p, b, hipMemcpyHostToDevice, s));
    fock_build(s, fock_args(c, kInitCopy, 0.0));
    MHIP_CHECK(hipMemcpyAsync(f, c->f.ptr, b, hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipStreamSynchronize(s));
  });
}

int mhip_pulay_sequence(mhip_ctx* c, int ncalls, double* f_seq, const double* pa_seq, double* pl_seq) {
  return guarded([&] {
    MHIP_REQUIRE(c && c->have_system && ncalls >= 0 && (ncalls == 0 || (f_seq && pa_seq && pl_seq)),
                 "mhip_pulay_sequence: bad arguments");
    const long long mp = c->mpack;
    const size_t b = sizeof(double) * mp;
    hipStream_t s = c->stream;
    PulayState st;
    for (int k = 0; k < ncalls; ++k) {
      MHIP_CHECK(hipMemcpyAsync(c->f.ptr, f_seq + (size_t)k * mp, b, hipMemcpyHostToDevice, s));
      MHIP_CHECK(hipMemcpyAsync(c->pa.ptr, pa_seq + (size_t)k * mp, b, hipMemcpyHostToDevice, s));
      double pl = 0.0;
      pulay_step(c, st, &pl);
      MHIP_CHECK(hipMemcpyAsync(f_seq + (size_t)k * mp, c->f.ptr, b, hipMemcpyDeviceToHost, s));
      MHIP_CHECK(hipStreamSynchronize(s));
      pl_seq[k] = pl;
    }
  });
}

int mhip_helect(mhip_ctx* c, int norbs, const double* p, const double* h, const double* f,
                double* e) {
  return guarded([&] {
    MHIP_REQUIRE(c && c->have_system && norbs == c->norbs && p && h && f && e, "mhip_helect: bad arguments");
    const size_t b = sizeof(double) * c->mpack;
    hipStream_t s = c->stream;
    MHIP_CHECK(hipMemcpyAsync(c->pa.ptr, p, b, hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(c->fcopy.ptr, h, b, hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(c->f.ptr, f, b, hipMemcpyHostToDevice, s));
    helect(s, norbs, c->pa.as<double>(), c->fcopy.as<double>(), c->f.as<double>(), 1.0,
           c->partial.as<double>(), c->scal.as<double>() + 16);
    *e = read_scalar(c, c->scal.as<double>() + 16);
  });
}

int mhip_eigen(mhip_ctx* c, double* cvec, double* fpacked, double* eig, int ndim) {
  return guarded([&] {
    MHIP_REQUIRE(c && c->have_system && ndim == c->norbs && cvec && fpacked && eig, "mhip_eigen: bad arguments");
    const size_t b = sizeof(double) * c->mpack;
    hipStream_t s = c->stream;
    MHIP_CHECK(hipMemcpyAsync(c->fcopy.ptr, fpacked, b, hipMemcpyHostToDevice, s));
    c->have_c = false;
    eigen_packed(c, c->fcopy.as<double>(), true);
    // the reference perturbs its argument in place
    MHIP_CHECK(hipMemcpyAsync(fpacked, c->fcopy.ptr, b, hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipMemcpyAsync(eig, c->eig.ptr, sizeof(double) * ndim, hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipMemcpy2DAsync(cvec, sizeof(double) * ndim, c->C.ptr, sizeof(double) * c->npad,
                                sizeof(double) * ndim, ndim, hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipStreamSynchronize(s));
  });
}

int mhip_pseudo_diag(mhip_ctx* c, const double* fao, double* vector, int nocc, const double* eig,
                     int mdim, int n, long long* nrot) {
  return guarded([&] {
    MHIP_REQUIRE(c && c->have_system && n == c->norbs && mdim == n && fao && vector && eig,
                 "mhip_pseudo_diag: bad arguments");
    MHIP_REQUIRE(nocc > 0 && nocc < n, "mhip_pseudo_diag: nocc out of range");
    hipStream_t s = c->stream;
    MHIP_CHECK(hipMemcpyAsync(c->fcopy.ptr, fao, sizeof(double) * c->mpack, hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpy2DAsync(c->C.ptr, sizeof(double) * c->npad, vector, sizeof(double) * n,
                                sizeof(double) * n, n, hipMemcpyHostToDevice, s));
    MHIP_CHECK(hipMemcpyAsync(c->eig.ptr, eig, sizeof(double) * n, hipMemcpyHostToDevice, s));
    long long nr = 0;
    pseudo_diag_device(s, c->pseudo, n, c->npad, c->fcopy.as<double>(), c->C.as<double>(), nocc,
                       c->eig.as<double>(), &nr);
    MHIP_CHECK(hipMemcpy2DAsync(vector, sizeof(double) * n, c->C.ptr, sizeof(double) * c->npad,
                                sizeof(double) * n, n, hipMemcpyDeviceToHost, s));
    MHIP_CHECK(hipStreamSynchronize(s));
    if (nrot) *nrot = nr;
    c->have_c = false;
  });
}

int mhip_scf(mhip_ctx* c, const mhip_scf_opts* opts, mhip_scf_result* res, double* p, double* f,
             double* cvec, double* eig) {
  return guarded([&] {
    MHIP_REQUIRE(c && opts, "mhip_scf: bad arguments");
    MHIP_CHECK(hipSetDevice(c->device));
    run_scf(c, opts, res, p, f, cvec, eig);
  });
}

int mhip_guess_water(mhip_ctx* c, mhip_mlp* mlp, const mhip_guess_model* model, const int* nat,
                     const double* xyz, int nw, const int* waters, int npairs, const int* pair_a,
                     const int* pair_b, float* x_out, float* y_out) {
  return guarded([&] {
    MHIP_REQUIRE(c && mlp && model && nat && xyz && waters && pair_a && pair_b,
                 "mhip_guess_water: null argument");
    MHIP_REQUIRE(model->x_mean && model->x_std && model->y_mean && model->y_std,
                 "mhip_guess_water: model normalization arrays missing");
    MHIP_REQUIRE(c->have_system, "mhip_guess_water: no system uploaded");
    MHIP_CHECK(hipSetDevice(c->device));
    const int n = c->numat;
    MHIP_REQUIRE(nw > 0 && 3 * nw == n, "mhip_guess_water: the system is not nw water molecules");
    MHIP_REQUIRE(npairs > 0, "mhip_guess_water: empty pair list");
    const int nl = mhip::mlp_layers(mlp);
    MHIP_REQUIRE(mhip::mlp_dim(mlp, 0) == mhip::kGuessNin && mhip::mlp_dim(mlp, nl) == mhip::kGuessNout,
                 "mhip_guess_water: network is not 96 -> ... -> 16");
    std::vector<int> mol(n, -1);
    for (int k = 0; k < nw; ++k)
      for (int q = 0; q < 3; ++q) {
        const int at = waters[3 * k + q];
        MHIP_REQUIRE(at >= 0 && at < n && mol[at] < 0, "mhip_guess_water: waters is not a partition of the atoms");
        mol[at] = k;
        const bool oxy = q == 0;
        MHIP_REQUIRE(nat[at] == (oxy ? 8 : 1) && c->no_h[at] == (oxy ? 4 : 1),
                     "mhip_guess_water: each water must be O (4 orbitals), H, H (1 orbital each)");
      }
    for (int m = 0; m < npairs; ++m)
      MHIP_REQUIRE(pair_b[m] >= 0 && pair_b[m] <= pair_a[m] && pair_a[m] < n,
                   "mhip_guess_water: pairs must satisfy 0 <= b <= a < numat");
    mhip::GuessHost in{n,           nw,          npairs,        nat,           xyz,
                       mol.data(),  waters,      pair_a,        pair_b,        c->nf0_h.data(),
                       c->no_h.data(), model->x_mean, model->x_std, model->y_mean, model->y_std};
    c->pguess.ensure(sizeof(double) * c->mpack);
    mhip::guess_water(c->stream, c->gw, mlp, in, c->pguess.as<double>(), c->mpack, x_out, y_out);
    c->have_guess = true;
  });
}

int mhip_guess_density(mhip_ctx* c, double* p) {
  return guarded([&] {
    MHIP_REQUIRE(c && p && c->have_guess, "mhip_guess_density: no guess on the device");
    MHIP_CHECK(hipSetDevice(c->device));
    MHIP_CHECK(hipMemcpyAsync(p, c->pguess.ptr, sizeof(double) * c->mpack, hipMemcpyDeviceToHost, c->stream));
    MHIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int mhip_last_sweeps(mhip_ctx* c) { return c ? c->last_sweeps : 0; }

// Event timing of the block-Jacobi kernels on the context stream: enable (1) /
// disable (0) and reset; read back accumulated milliseconds and launch counts.
const char* mhip_kernel_class_name(int cls) {
  static const char* names[K_NCLASS] = {"k_gemm_mfma (FP64 GEMM)", "k_gemm_split_bf16 (split-BF16 GEMM)",
                                        "k_syrk_packed_mfma (density)", "k_fock_rows+k_fock_cols (fock2)"};
  return (cls >= 0 && cls < K_NCLASS) ? names[cls] : "";
}

int mhip_kernel_stats(mhip_ctx* c, int enable, int max_classes, double* ms, long long* launches,
                      double* flops, double* bytes) {
  return guarded([&] {
    MHIP_REQUIRE(c != nullptr, "mhip_kernel_stats: null context");
    MHIP_CHECK(hipSetDevice(c->device));
    MHIP_CHECK(hipStreamSynchronize(c->stream));
    if (enable == 0 || enable == 1) {
      g_ktimer.recs.clear();
      g_ktimer.used = 0;
      g_ktimer.on = enable == 1;
      return;
    }
    const int nc = std::min(max_classes, (int)K_NCLASS);
    for (int i = 0; i < nc; ++i) {
      if (ms) ms[i] = 0.0;
      if (launches) launches[i] = 0;
      if (flops) flops[i] = 0.0;
      if (bytes) bytes[i] = 0.0;
    }
    for (const auto& r : g_ktimer.recs) {
      if (r.cls >= nc) continue;
      float e = 0.f;
      MHIP_CHECK(hipEventSynchronize(r.b));
      MHIP_CHECK(hipEventElapsedTime(&e, r.a, r.b));
      if (ms) ms[r.cls] += e;
      if (launches) launches[r.cls] += 1;
      if (flops) flops[r.cls] += r.flops;
      if (bytes) bytes[r.cls] += r.bytes;
    }
  });
}

int mhip_kernel_timing(mhip_ctx* c, int enable, double* solve_ms, long long* n_solve,
                       double* apply_ms, long long* n_apply, double* applyc_ms,
                       long long* n_applyc, long long* active_tiles, long long* active_pairs,
                       long long* rotations) {
  if (!c) return 1;
  // both plans: the in-loop one and the grouped one of the exit diagonalization
  const mhip::JacobiPlan& a = c->plan;
  const mhip::JacobiPlan& b = c->plan_split;
  if (active_tiles) *active_tiles = a.active_tiles + b.active_tiles;
  if (active_pairs) *active_pairs = a.active_pairs + b.active_pairs;
  if (rotations) *rotations = a.active_rotations + b.active_rotations;
  if (solve_ms) *solve_ms = a.t_solve_ms + b.t_solve_ms;
  if (n_solve) *n_solve = a.n_solve + b.n_solve;
  if (apply_ms) *apply_ms = a.t_apply_ms + b.t_apply_ms;
  if (n_apply) *n_apply = a.n_apply + b.n_apply;
  if (applyc_ms) *applyc_ms = a.t_applyc_ms + b.t_applyc_ms;
  if (n_applyc) *n_applyc = a.n_applyc + b.n_applyc;
  if (enable >= 0) {
    for (mhip::JacobiPlan* p : {&c->plan, &c->plan_split}) {
      p->timing = enable != 0;
      p->t_solve_ms = p->t_apply_ms = p->t_applyc_ms = 0.0;
      p->n_solve = p->n_apply = p->n_applyc = 0;
      p->active_tiles = p->active_pairs = p->active_rotations = 0;
    }
  }
  return 0;
}

// Diagnostics/tuning (not part of the reference boundary): kernel family
// (1 = FP64 MFMA, 0 = vector FMA), inner Jacobi sweeps, in-loop tolerance.
int mhip_set_tuning(mhip_ctx* c, int kernel_variant, int max_inner, double tol_iter) {
  if (kernel_variant >= 0) mhip::g_kernel_variant = kernel_variant;
  if (c && max_inner > 0) c->plan.max_inner = max_inner;
  if (c && tol_iter > 0) c->tol_iter = tol_iter;
  return 0;
}

// Pair-solve schedule of the block-Jacobi eigensolver, process-wide:
// 1 = cross rounds + one intra-block pass per sweep, 0 = full inner sweeps.
int mhip_set_jacobi_schedule(int schedule) {
  const int prev = mhip::g_jacobi_schedule;
  if (schedule == 0 || schedule == 1) mhip::g_jacobi_schedule = schedule;
  return prev;
}

}  // extern "C"
