// FP64 matrix-core (MFMA) helpers for gfx950.
//
// v_mfma_f64_16x16x4f64 lane layout (64 lanes, l = lane id):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l holds D[l / 16 + 4 * t][l % 16] in acc[t], t = 0..3
//              (rows interleaved by 4, measured on gfx950 with structured
//              operands; the GEMM parity test re-checks it on every run)
#pragma once
#include <hip/hip_runtime.h>

namespace mhip {

typedef double d4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4_t mfma16x16x4(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 64x64 output tile = sum_k A(r,k) B(k,c) over k in [0,K) with A, B read by
// functors from LDS.  Four waves; wave w owns the 32x32 quadrant
// (rows 32*(w/2), cols 32*(w%2)) as 2x2 MFMA blocks; acc[bi*2+bj].
template <class FA, class FB>
__device__ __forceinline__ void mfma_tile64(const FA& A, const FB& B, int K, d4_t acc[4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qr = 32 * (w >> 1), qc = 32 * (w & 1);
  const int li = lane & 15, lk = lane >> 4;
  for (int k = 0; k < K; k += 4) {
    const double a0 = A(qr + li, k + lk), a1 = A(qr + 16 + li, k + lk);
    const double b0 = B(k + lk, qc + li), b1 = B(k + lk, qc + 16 + li);
    acc[0] = mfma16x16x4(a0, b0, acc[0]);
    acc[1] = mfma16x16x4(a0, b1, acc[1]);
    acc[2] = mfma16x16x4(a1, b0, acc[2]);
    acc[3] = mfma16x16x4(a1, b1, acc[3]);
  }
}

// Visit the 16 outputs a lane owns after mfma_tile64: fn(row, col, value).
template <class F>
__device__ __forceinline__ void mfma_tile64_store(const d4_t acc[4], const F& fn) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qr = 32 * (w >> 1), qc = 32 * (w & 1);
  const int col = lane & 15, rb = lane >> 4;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int t = 0; t < 4; ++t) fn(qr + 16 * bi + rb + 4 * t, qc + 16 * bj + col, acc[bi * 2 + bj][t]);
}

// The same 64x64 tile with eight waves: wave w owns the 32x16 strip at rows
// 32*(w/4), cols 16*(w%4) as two MFMA blocks stacked vertically; acc[bi].
template <class FA, class FB>
__device__ __forceinline__ void mfma_tile64_w8(const FA& A, const FB& B, int K, d4_t acc[2]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qr = 32 * (w >> 2), qc = 16 * (w & 3);
  const int li = lane & 15, lk = lane >> 4;
  for (int k = 0; k < K; k += 4) {
    const double a0 = A(qr + li, k + lk), a1 = A(qr + 16 + li, k + lk);
    const double b0 = B(k + lk, qc + li);
    acc[0] = mfma16x16x4(a0, b0, acc[0]);
    acc[1] = mfma16x16x4(a1, b0, acc[1]);
  }
}

template <class F>
__device__ __forceinline__ void mfma_tile64_w8_store(const d4_t acc[2], const F& fn) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qr = 32 * (w >> 2), qc = 16 * (w & 3);
  const int col = lane & 15, rb = lane >> 4;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int t = 0; t < 4; ++t) fn(qr + 16 * bi + rb + 4 * t, qc + col, acc[bi][t]);
}

}  // namespace mhip
