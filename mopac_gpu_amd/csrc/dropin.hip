// Reference-boundary drop-ins: the C signatures the reference's Fortran binds
// (src/gpu/cuda_wrappers.cu), served by this library's kernels on the device
// mhip_setDevice selected (device 0 by default) with grow-only device
// buffers.  Host arrays in, host arrays out, as there.
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "k_eigen.h"
#include "k_pseudo.h"

namespace mhip {
namespace {

struct Implicit {
  bool init = false;
  int device = 0;
  hipStream_t s = nullptr;
  DevBuf a, b, c, d, cols, rot, cnt;
  JacobiPlan plan;
  double* hs = nullptr;
  DevBuf scal;
};
Implicit g_imp;
std::mutex g_mu;
int g_selected_device = 0;    // set by mhip_setDevice
int g_pair[2] = {0, 1};       // mhip_set_mozyme_gpu_pair (recorded; one GPU is used)

// Release every buffer of the implicit context (caller holds g_mu).
void release_implicit() {
  if (!g_imp.init) return;
  (void)hipSetDevice(g_imp.device);
  for (DevBuf* b : {&g_imp.a, &g_imp.b, &g_imp.c, &g_imp.d, &g_imp.cols, &g_imp.rot, &g_imp.cnt,
                    &g_imp.scal})
    b->release();
  g_imp.plan.release();
  release_split_scratch(g_imp.s);
  release_splitk_scratch(g_imp.s);
  if (g_imp.hs) (void)hipHostFree(g_imp.hs);
  (void)hipStreamDestroy(g_imp.s);
  g_imp = Implicit();
}

// The implicit context on the selected device; a new selection releases the
// buffers held on the old one first.
Implicit& imp() {
  if (g_imp.init && g_imp.device != g_selected_device) release_implicit();
  if (!g_imp.init) {
    MHIP_CHECK(hipSetDevice(g_selected_device));
    g_imp.device = g_selected_device;
    MHIP_CHECK(hipStreamCreateWithFlags(&g_imp.s, hipStreamNonBlocking));
    MHIP_CHECK(hipHostMalloc((void**)&g_imp.hs, sizeof(double) * 64, hipHostMallocDefault));
    g_imp.scal.ensure(sizeof(double) * 64);
    g_imp.init = true;
  } else {
    MHIP_CHECK(hipSetDevice(g_imp.device));
  }
  return g_imp;
}

void h2d2(DevBuf& b, const double* src, int rows, int cols, int ld, int ldd, hipStream_t s) {
  b.ensure(sizeof(double) * (size_t)ldd * cols + 8);
  MHIP_CHECK(hipMemcpy2DAsync(b.ptr, sizeof(double) * ldd, src, sizeof(double) * ld,
                              sizeof(double) * rows, cols, hipMemcpyHostToDevice, s));
}

void report(const HipError& e) {
  set_error(e.msg);
  std::fprintf(stderr, "libmopachip: %s\n", e.msg.c_str());
}

}  // namespace
}  // namespace mhip

using namespace mhip;

extern "C" {

}  // extern "C"

static void call_gemm_impl(bool split, char tra, char trb, int m, int n, int k, double alpha, const double* a,
                    int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    Implicit& I = imp();
    const bool ta = (tra == 'T' || tra == 't'), tb = (trb == 'T' || trb == 't');
    const int ar = ta ? k : m, ac = ta ? m : k;
    const int br = tb ? n : k, bc = tb ? k : n;
    h2d2(I.a, a, ar, ac, lda, ar, I.s);
    h2d2(I.b, b, br, bc, ldb, br, I.s);
    h2d2(I.c, c, m, n, ldc, m, I.s);
    if (split)
      gemm_split_bf16(I.s, tra, trb, m, n, k, alpha, I.a.as<double>(), ar, I.b.as<double>(), br,
                      beta, I.c.as<double>(), m);
    else
      gemm(I.s, tra, trb, m, n, k, alpha, I.a.as<double>(), ar, I.b.as<double>(), br, beta,
           I.c.as<double>(), m);
    MHIP_CHECK(hipMemcpy2DAsync(c, sizeof(double) * ldc, I.c.ptr, sizeof(double) * m,
                                sizeof(double) * m, n, hipMemcpyDeviceToHost, I.s));
    MHIP_CHECK(hipStreamSynchronize(I.s));
  } catch (const HipError& e) {
    report(e);
  }
}

extern "C" {

void mhip_call_gemm(char tra, char trb, int m, int n, int k, double alpha, const double* a,
                    int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  call_gemm_impl(false, tra, trb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void mhip_gemm_split_bf16(char tra, char trb, int m, int n, int k, double alpha, const double* a,
                          int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  call_gemm_impl(true, tra, trb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void mhip_call_syrk(char uplo, char tra, int n, int k, double alpha, const double* a, int lda,
                    double beta, double* c, int ldc) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    Implicit& I = imp();
    const bool ta = (tra == 'T' || tra == 't');
    const int ar = ta ? k : n, ac = ta ? n : k;
    h2d2(I.a, a, ar, ac, lda, ar, I.s);
    I.d.ensure(sizeof(double) * (size_t)n * n + 8);
    // D = op(A) op(A)^T, symmetric: lower-triangle tiles, then mirrored
    if (ta) gemm_sym(I.s, 'T', 'N', n, k, alpha, I.a.as<double>(), ar, I.a.as<double>(), ar, 0.0, I.d.as<double>(), n);
    else gemm_sym(I.s, 'N', 'T', n, k, alpha, I.a.as<double>(), ar, I.a.as<double>(), ar, 0.0, I.d.as<double>(), n);
    std::vector<double> d((size_t)n * n);
    MHIP_CHECK(hipMemcpyAsync(d.data(), I.d.ptr, sizeof(double) * d.size(), hipMemcpyDeviceToHost, I.s));
    MHIP_CHECK(hipStreamSynchronize(I.s));
    const bool up = (uplo == 'U' || uplo == 'u');
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        if (up ? (i > j) : (i < j)) continue;
        double& dst = c[(size_t)i + (size_t)j * ldc];
        dst = d[(size_t)i + (size_t)j * n] + (beta == 0.0 ? 0.0 : beta * dst);
      }
  } catch (const HipError& e) {
    report(e);
  }
}

void mhip_dsyevd(int n, double* a, int lda, double* w, int* info) {
  std::lock_guard<std::mutex> lk(g_mu);
  *info = 0;
  try {
    Implicit& I = imp();
    I.plan.build(n);
    const int np = I.plan.npad;
    // symmetric from the upper triangle, padded with decoupled large entries
    std::vector<double> full((size_t)np * np, 0.0);
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i) {
        double v;
        if (i < n && j < n) v = (i <= j) ? a[(size_t)i + (size_t)j * lda] : a[(size_t)j + (size_t)i * lda];
        else v = (i == j) ? 1.0e4 + i : 0.0;
        full[(size_t)i + (size_t)j * np] = v;
      }
    I.a.ensure(sizeof(double) * full.size());
    I.c.ensure(sizeof(double) * full.size());
    I.d.ensure(sizeof(double) * full.size());
    I.b.ensure(sizeof(double) * np);
    MHIP_CHECK(hipMemcpyAsync(I.a.ptr, full.data(), sizeof(double) * full.size(), hipMemcpyHostToDevice, I.s));
    jacobi_set_identity(I.s, np, I.c.as<double>(), np);
    int sweeps = 0;
    if (jacobi_sweeps(I.s, I.plan, I.a.as<double>(), I.c.as<double>(), 1.0e-12, 60, I.hs,
                      I.scal.as<double>(), &sweeps) != 0)
      *info = 1;
    jacobi_sort(I.s, I.plan, I.a.as<double>(), I.c.as<double>(), I.d.as<double>(), I.b.as<double>());
    MHIP_CHECK(hipMemcpy2DAsync(a, sizeof(double) * lda, I.d.ptr, sizeof(double) * np,
                                sizeof(double) * n, n, hipMemcpyDeviceToHost, I.s));
    MHIP_CHECK(hipMemcpyAsync(w, I.b.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, I.s));
    MHIP_CHECK(hipStreamSynchronize(I.s));
  } catch (const HipError& e) {
    report(e);
    *info = -1;
  }
}

void mhip_call_rot(const double* fmo, const double* eig, double* v, const double* ci0,
                   const double* ca0, int nocc, int lumo, int n, double bigeps, double tiny) {
  (void)ci0;
  (void)ca0;
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    Implicit& I = imp();
    // ordered list exactly as call_rot_cuda_gpu walks it (occupied outer)
    std::vector<int2> cols;
    std::vector<double2> rot;
    long long ij = 0;
    for (int i = 0; i < nocc; ++i)
      for (int j = lumo - 1; j < n; ++j) {
        ij += 1;
        const double x = fmo[ij - 1];
        if (std::fabs(x) < tiny) continue;
        const double d = eig[i] - eig[j];
        if (std::fabs(x / d) < bigeps) continue;
        const double e = std::copysign(std::sqrt(4.0 * x * x + d * d), d);
        const double alpha = std::sqrt(0.5 * (1.0 + d / e));
        const double beta = -std::copysign(std::sqrt(1.0 - alpha * alpha), x);
        cols.push_back(make_int2(i, j));
        rot.push_back(make_double2(alpha, beta));
      }
    const long long nr = (long long)cols.size();
    I.cols.ensure(sizeof(int2) * (nr + 1));
    I.rot.ensure(sizeof(double2) * (nr + 1));
    I.cnt.ensure(sizeof(long long));
    h2d2(I.c, v, n, n, n, n, I.s);
    if (nr) {
      MHIP_CHECK(hipMemcpyAsync(I.cols.ptr, cols.data(), sizeof(int2) * nr, hipMemcpyHostToDevice, I.s));
      MHIP_CHECK(hipMemcpyAsync(I.rot.ptr, rot.data(), sizeof(double2) * nr, hipMemcpyHostToDevice, I.s));
    }
    MHIP_CHECK(hipMemcpyAsync(I.cnt.ptr, &nr, sizeof(long long), hipMemcpyHostToDevice, I.s));
    apply_rotation_list(I.s, n, n, I.c.as<double>(), I.cnt.as<long long>(), I.cols.as<int2>(),
                        I.rot.as<double2>());
    MHIP_CHECK(hipMemcpyAsync(v, I.c.ptr, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, I.s));
    MHIP_CHECK(hipStreamSynchronize(I.s));
  } catch (const HipError& e) {
    report(e);
  }
}

void mhip_destroy_resources(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  release_implicit();
}

void mhip_getGPUInfo(bool* hasGpu, bool hasDouble[6], int* nDevices, char name[6][256], int name_size[6],
                     size_t totalMem[6], int clockRate[6], int major[6], int minor[6]) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    if (hasGpu) *hasGpu = false;
    if (nDevices) *nDevices = 0;
    return;
  }
  const int nd = count > 6 ? 6 : count;
  if (hasGpu) *hasGpu = true;
  if (nDevices) *nDevices = nd;
  for (int i = 0; i < 6; ++i) {
    hipDeviceProp_t prop{};
    const bool ok = i < nd && hipGetDeviceProperties(&prop, i) == hipSuccess;
    if (ok) {
      // the marketing name when the driver has one, else the gfx target
      std::strncpy(name[i], prop.name[0] ? prop.name : prop.gcnArchName, 255);
      name[i][255] = '\0';
    } else {
      name[i][0] = '\0';
    }
    name_size[i] = ok ? (int)std::strlen(name[i]) : 0;
    totalMem[i] = ok ? prop.totalGlobalMem : 0;
    clockRate[i] = ok ? prop.clockRate : 0;
    major[i] = ok ? prop.major : 0;
    minor[i] = ok ? prop.minor : 0;
    hasDouble[i] = ok;  // every CDNA device has native FP64
  }
}

void mhip_setDevice(int idevice, bool* stat) {
  std::lock_guard<std::mutex> lk(g_mu);
  int count = 0;
  const bool ok = hipGetDeviceCount(&count) == hipSuccess && idevice >= 0 && idevice < count &&
                  hipSetDevice(idevice) == hipSuccess;
  if (ok) g_selected_device = idevice;
  if (stat) *stat = ok;
}

void mhip_set_mozyme_gpu_pair(int dev0, int dev1) {
  std::lock_guard<std::mutex> lk(g_mu);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return;
  if (dev0 >= 0 && dev0 < count) g_pair[0] = dev0;
  if (dev1 >= 0 && dev1 < count) g_pair[1] = dev1;
}

int mhip_selected_device(void) { return g_selected_device; }

}  // extern "C"
