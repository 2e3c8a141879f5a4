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

// ---- per-class launch timing (HIP events on the launching stream) --------
// Off unless mhip_kernel_stats switches it on; each timed launch records an
// event pair and its algorithmic flops and HBM bytes.
enum KClass {
  K_GEMM_F64 = 0,    // gemm on the FP64 MFMA tiles: 2 m n k flops
  K_GEMM_SPLIT,      // gemm_split_bf16: 2 m n k (FP64-equivalent) flops
  K_SYRK_F64,        // packed density: n (n + 1) k flops
  K_FOCK,            // fock_build (row + column pass): 8 (n2elec + 4 mpack) bytes
  K_NCLASS
};
struct KTimer {
  bool on = false;
  struct Rec {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t next() {
    if (used == pool.size()) {
      hipEvent_t e;
      MHIP_CHECK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
};
extern KTimer g_ktimer;
// Wrap one launch (or a short fixed sequence) of class cls.
template <class F>
inline void timed(hipStream_t s, int cls, double flops, double bytes, F&& launch) {
  if (!g_ktimer.on) {
    launch();
    return;
  }
  hipEvent_t a = g_ktimer.next(), b = g_ktimer.next();
  MHIP_CHECK(hipEventRecord(a, s));
  launch();
  MHIP_CHECK(hipEventRecord(b, s));
  g_ktimer.recs.push_back({cls, a, b, flops, bytes});
}

// ---- dense kernels (k_dense.hip) ------------------------------------------
// 1: FP64 MFMA tiles (default), 0: vector-FMA tiles (kept as a cross-check).
extern int g_kernel_variant;
// C = alpha * op(A) * op(B) + beta * C, column-major FP64.
void gemm(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double* a,
          int lda, const double* b, int ldb, double beta, double* c, int ldc);
// gemm for an m x m result known to be symmetric (a Gram matrix A^T A, or
// A^T B with A^T B = B^T A): only the tiles on and below the diagonal are
// multiplied, and the upper triangle is then copied from the lower one.  The
// input C (beta != 0) is read on the lower triangle only.
void gemm_sym(hipStream_t s, char ta, char tb, int m, int k, double alpha, const double* a, int lda,
              const double* b, int ldb, double beta, double* c, int ldc);
// Same contract as gemm, computed on the BF16 matrix cores from a three-way
// BF16 split of each FP64 operand (six cross products, FP32 accumulation in
// 32-term chunks summed in FP64; k_split.hip).
void gemm_split_bf16(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                     const double* a, int lda, const double* b, int ldb, double beta, double* c,
                     int ldc);
// Frees the BF16 plane buffers gemm_split_bf16 keeps per (device, stream):
// those of stream s, or all of them when s is null.
void release_split_scratch(hipStream_t s);
// Grow-only scratch for the K-slice parts of a split GEMM on stream s, and
// c(m x n, ldc) += the nparts parts (m x n each, in order) it holds.
double* splitk_scratch(hipStream_t s, size_t bytes);
void add_parts(hipStream_t s, int m, int n, int nparts, const double* part, double* c, int ldc);
// Frees the split-K GEMM scratch of stream s (all streams when s is null).
void release_splitk_scratch(hipStream_t s);
// Packed lower triangle p (mpack) = occ * A(:,0:k) A(:,0:k)^T over the first n rows.
void syrk_packed(hipStream_t s, int n, int k, double occ, const double* a, int lda, double* p);
// Split-BF16 variant (gemm_split_bf16 into a dense n x n scratch, then packed).
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
