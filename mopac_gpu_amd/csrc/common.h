// Shared definitions for libmopachip (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "mopac_hip.h"

namespace mhip {

// Thread-local last-error text, returned by mhip_last_error().
void set_error(const std::string& msg);
const std::string& get_error();

struct HipError {
  std::string msg;
};

#define MHIP_CHECK(expr)                                                                  \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) {                                                               \
      throw ::mhip::HipError{std::string(#expr) + " failed: " + hipGetErrorString(_e) +  \
                             " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"};     \
    }                                                                                     \
  } while (0)

#define MHIP_LAUNCH_CHECK() MHIP_CHECK(hipGetLastError())

#define MHIP_REQUIRE(cond, text)                  \
  do {                                            \
    if (!(cond)) throw ::mhip::HipError{(text)};  \
  } while (0)

// A grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void ensure(size_t need) {
    if (need <= bytes) return;
    if (ptr) MHIP_CHECK(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    if (need == 0) return;
    MHIP_CHECK(hipMalloc(&ptr, need));
    bytes = need;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(ptr);
  }
};

inline long long mpack_of(long long n) { return n * (n + 1) / 2; }

// ---- dense kernels (k_dense.hip) ------------------------------------------
// 1: FP64 MFMA tiles (default), 0: vector-FMA tiles (kept as a cross-check).
extern int g_kernel_variant;
// C = alpha * op(A) * op(B) + beta * C, column-major FP64.
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double* a,
          int lda, const double* b, int ldb, double beta, double* c, int ldc);
// Packed lower triangle p (mpack) = occ * A(:,0:k) A(:,0:k)^T over the first n rows.
void syrk_packed(hipStream_t s, int n, int k, double occ, const double* a, int lda, double* p);
// Split-BF16 variant: A = hi + lo (both bf16), products accumulated in FP32 on
// MFMA, with the lo*lo term dropped and an FP64 correction of the diagonal.
void syrk_packed_split_bf16(hipStream_t s, int n, int k, double occ, const double* a, int lda,
                            double* p, void* scratch);
size_t syrk_split_scratch_bytes(int n, int k);
// dense (ld) <-> packed lower
void unpack_sym(hipStream_t s, int n, const double* p, double* a, int lda, int npad,
                double pad_base);
void pack_lower(hipStream_t s, int n, const double* a, int lda, double* p);
// Packed antisymmetric commutator out(r,c) = X(c,r) - X(r,c) for r >= c.
void pack_antisym(hipStream_t s, int n, const double* x, int ldx, double* out);
// y[i] = sum_k coef[k] * x[k*stride + i]
void lincomb(hipStream_t s, long long len, int nk, const double* coef_dev, const double* x,
             long long stride, double* y);
// Fill with a scalar.
void fill(hipStream_t s, double* p, long long n, double v);

// ---- reductions ------------------------------------------------------------
// Deterministic dot products: out[j] = <a_j, b> for j < nvec, a_j = a + j*stride.
void dots(hipStream_t s, long long len, int nvec, const double* a, long long stride,
          const double* b, double* out, double* partial);
size_t dots_partial_doubles(long long len, int nvec);

}  // namespace mhip
