// Small MLP for the learned starting density: BF16 weights and layer inputs,
// FP32 accumulation, tanh hidden activations, linear output.  Forward,
// backward (mean-squared error), and Adam, all on the device.
//
// Layer l: Z = bf16(X_l) * bf16(W_l) + b_l (FP32), X_{l+1} = tanh(Z) (hidden),
// Y = Z (last).  Master weights stay FP32; the BF16 copies are refreshed after
// every optimizer step.  Backward treats the BF16 rounding of activations as
// the identity (straight-through), the usual mixed-precision convention.
#include <vector>

#include "common.h"

namespace mhip {
namespace {

__device__ __forceinline__ unsigned short f2bf(float x) {
  // round to nearest even
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

constexpr int T = 32;  // tile edge

// C[M x N] = op(A) * op(B) (+bias), FP32 accumulate.  A, B either BF16
// (unsigned short) or FP32.  ta/tb select transposes of row-major storage:
//   A row-major M x K (ta = 0) or K x M (ta = 1); B row-major K x N or N x K.
template <class TA_, class TB_, bool TRA, bool TRB>
__global__ __launch_bounds__(256) void k_mm(int M, int N, int K, const TA_* __restrict__ A,
                                            const TB_* __restrict__ B, const float* __restrict__ bias,
                                            float* __restrict__ C, int act, float beta) {
  __shared__ float As[T][T + 1];
  __shared__ float Bs[T][T + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int row0 = blockIdx.y * T, col0 = blockIdx.x * T;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  auto lda = [&](int m, int k) -> float {
    if (m >= M || k >= K) return 0.f;
    size_t idx = TRA ? (size_t)k * M + m : (size_t)m * K + k;
    if constexpr (sizeof(TA_) == 2) return bf2f(A[idx]);
    else return (float)A[idx];
  };
  auto ldb = [&](int k, int n) -> float {
    if (k >= K || n >= N) return 0.f;
    size_t idx = TRB ? (size_t)n * K + k : (size_t)k * N + n;
    if constexpr (sizeof(TB_) == 2) return bf2f(B[idx]);
    else return (float)B[idx];
  };
  for (int k0 = 0; k0 < K; k0 += T) {
    for (int r = ty; r < T; r += 8) {
      As[r][tx] = lda(row0 + r, k0 + tx);
      Bs[r][tx] = ldb(k0 + r, col0 + tx);
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < T; ++kk) {
      const float b = Bs[kk][tx];
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = fmaf(As[ty + 8 * i][kk], b, acc[i]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = row0 + ty + 8 * i, c = col0 + tx;
    if (r >= M || c >= N) continue;
    float v = acc[i] + (bias ? bias[c] : 0.f);
    if (act == 1) v = tanhf(v);
    float* dst = C + (size_t)r * N + c;
    *dst = (beta == 0.f) ? v : v + beta * *dst;
  }
}

__global__ void k_to_bf16(long long n, const float* __restrict__ x, unsigned short* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// dY = (Y - T) * 2 / (B * nout); partial loss sums per block
__global__ __launch_bounds__(256) void k_mse(long long n, float scale, const float* __restrict__ y,
                                             const float* __restrict__ t, float* __restrict__ dy,
                                             float* __restrict__ part) {
  __shared__ float red[256];
  float s = 0.f;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float d = y[i] - t[i];
    s += d * d;
    dy[i] = 2.f * d * scale;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_f(int n, const float* __restrict__ part, float scale, float* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = (float)(s * scale);
  }
}

// dZ = dX_next * (1 - tanh^2) where X_next = tanh(Z) is stored
__global__ void k_tanh_back(long long n, const float* __restrict__ xn, float* __restrict__ d) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float a = xn[i];
    d[i] *= (1.f - a * a);
  }
}

// db[c] = sum_r dZ[r][c]
__global__ void k_colsum(int rows, int cols, const float* __restrict__ dz, float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += dz[(size_t)r * cols + c];
  db[c] = s;
}

__global__ void k_adam(long long n, float* __restrict__ w, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace
}  // namespace mhip

using namespace mhip;

struct mhip_mlp {
  int nl = 0;
  std::vector<int> dims;
  std::vector<size_t> woff, boff;  // offsets into the flat parameter vector
  size_t np = 0;
  hipStream_t s = nullptr;
  DevBuf w32, wbf, grad, m, v;
  DevBuf acts;  // X_0..X_nl (FP32) for the current batch
  DevBuf xbf;   // BF16 copy of a layer input
  DevBuf dbuf0, dbuf1, tgt, part, loss;
  int batch_cap = 0;
  long long step = 0;
};

namespace {

void mlp_ensure_batch(mhip_mlp* q, int batch) {
  if (batch <= q->batch_cap) return;
  size_t tot = 0, mx = 0;
  for (int l = 0; l <= q->nl; ++l) {
    tot += (size_t)batch * q->dims[l];
    mx = std::max(mx, (size_t)q->dims[l]);
  }
  q->acts.ensure(sizeof(float) * tot);
  q->xbf.ensure(sizeof(unsigned short) * (size_t)batch * mx);
  q->dbuf0.ensure(sizeof(float) * (size_t)batch * mx);
  q->dbuf1.ensure(sizeof(float) * (size_t)batch * mx);
  q->tgt.ensure(sizeof(float) * (size_t)batch * q->dims[q->nl]);
  q->part.ensure(sizeof(float) * 1024);
  q->loss.ensure(sizeof(float) * 4);
  q->batch_cap = batch;
}

float* act_ptr(mhip_mlp* q, int batch, int l) {
  size_t off = 0;
  for (int i = 0; i < l; ++i) off += (size_t)batch * q->dims[i];
  return q->acts.as<float>() + off;
}

void refresh_bf16(mhip_mlp* q) {
  hipLaunchKernelGGL(k_to_bf16, dim3(std::min(4096, cdiv(q->np, 256))), dim3(256), 0, q->s,
                     (long long)q->np, q->w32.as<float>(), q->wbf.as<unsigned short>());
  MHIP_LAUNCH_CHECK();
}

void forward(mhip_mlp* q, int batch) {
  for (int l = 0; l < q->nl; ++l) {
    const int in = q->dims[l], out = q->dims[l + 1];
    float* X = act_ptr(q, batch, l);
    float* Y = act_ptr(q, batch, l + 1);
    hipLaunchKernelGGL(k_to_bf16, dim3(std::min(4096, cdiv((long long)batch * in, 256))), dim3(256), 0,
                       q->s, (long long)batch * in, X, q->xbf.as<unsigned short>());
    MHIP_LAUNCH_CHECK();
    const unsigned short* W = q->wbf.as<unsigned short>() + q->woff[l];
    const float* b = q->w32.as<float>() + q->boff[l];
    hipLaunchKernelGGL((k_mm<unsigned short, unsigned short, false, false>),
                       dim3(cdiv(out, T), cdiv(batch, T)), dim3(256), 0, q->s, batch, out, in,
                       q->xbf.as<unsigned short>(), W, b, Y, (l + 1 < q->nl) ? 1 : 0, 0.f);
    MHIP_LAUNCH_CHECK();
  }
}

// gradients of the mean squared error into q->grad; loss on the device
void backward(mhip_mlp* q, int batch) {
  const int nout = q->dims[q->nl];
  const long long ny = (long long)batch * nout;
  float* Y = act_ptr(q, batch, q->nl);
  const int pb = std::min(1024, cdiv(ny, 256));
  hipLaunchKernelGGL(k_mse, dim3(pb), dim3(256), 0, q->s, ny, 1.f / (float)ny, Y, q->tgt.as<float>(),
                     q->dbuf0.as<float>(), q->part.as<float>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_f, dim3(1), dim3(64), 0, q->s, pb, q->part.as<float>(), 1.f / (float)ny,
                     q->loss.as<float>());
  MHIP_LAUNCH_CHECK();
  float* dcur = q->dbuf0.as<float>();
  float* dnext = q->dbuf1.as<float>();
  for (int l = q->nl - 1; l >= 0; --l) {
    const int in = q->dims[l], out = q->dims[l + 1];
    float* X = act_ptr(q, batch, l);
    if (l + 1 < q->nl) {
      // hidden layer: through tanh
      hipLaunchKernelGGL(k_tanh_back, dim3(std::min(4096, cdiv((long long)batch * out, 256))), dim3(256),
                         0, q->s, (long long)batch * out, act_ptr(q, batch, l + 1), dcur);
      MHIP_LAUNCH_CHECK();
    }
    // dW = bf16(X)^T dZ  (in x out), db = colsum(dZ)
    hipLaunchKernelGGL(k_to_bf16, dim3(std::min(4096, cdiv((long long)batch * in, 256))), dim3(256), 0,
                       q->s, (long long)batch * in, X, q->xbf.as<unsigned short>());
    MHIP_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_mm<unsigned short, float, true, false>), dim3(cdiv(out, T), cdiv(in, T)),
                       dim3(256), 0, q->s, in, out, batch, q->xbf.as<unsigned short>(), dcur,
                       (const float*)nullptr, q->grad.as<float>() + q->woff[l], 0, 0.f);
    MHIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_colsum, dim3(cdiv(out, 256)), dim3(256), 0, q->s, batch, out, dcur,
                       q->grad.as<float>() + q->boff[l]);
    MHIP_LAUNCH_CHECK();
    if (l > 0) {
      // dX = dZ W^T (batch x in), W in BF16
      hipLaunchKernelGGL((k_mm<float, unsigned short, false, true>), dim3(cdiv(in, T), cdiv(batch, T)),
                         dim3(256), 0, q->s, batch, in, out, dcur,
                         q->wbf.as<unsigned short>() + q->woff[l], (const float*)nullptr, dnext, 0, 0.f);
      MHIP_LAUNCH_CHECK();
      std::swap(dcur, dnext);
    }
  }
}

template <class F>
int guard(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const HipError& e) {
    set_error(e.msg);
    return 1;
  }
}

}  // namespace

// Device-resident forward pass for callers in other files (the learned
// guess): they write the batch into mlp_input, then mlp_forward_resident
// runs the network on the network's own stream, waits for it and returns the
// device pointer of the outputs (batch x dims[nl], row-major).
namespace mhip {
float* mlp_input(mhip_mlp* q, int batch) {
  mlp_ensure_batch(q, batch);
  return act_ptr(q, batch, 0);
}
const float* mlp_forward_resident(mhip_mlp* q, int batch) {
  forward(q, batch);
  MHIP_CHECK(hipStreamSynchronize(q->s));
  return act_ptr(q, batch, q->nl);
}
int mlp_layers(const mhip_mlp* q) { return q->nl; }
int mlp_dim(const mhip_mlp* q, int l) { return q->dims[l]; }
}  // namespace mhip

extern "C" {

int mhip_mlp_create(int nlayers, const int* dims, mhip_mlp** out) {
  return guard([&] {
    MHIP_REQUIRE(nlayers > 0 && dims && out, "mhip_mlp_create: bad arguments");
    auto* q = new mhip_mlp();
    q->nl = nlayers;
    q->dims.assign(dims, dims + nlayers + 1);
    size_t off = 0;
    for (int l = 0; l < nlayers; ++l) {
      q->woff.push_back(off);
      off += (size_t)dims[l] * dims[l + 1];
      q->boff.push_back(off);
      off += dims[l + 1];
    }
    q->np = off;
    MHIP_CHECK(hipStreamCreateWithFlags(&q->s, hipStreamNonBlocking));
    q->w32.ensure(sizeof(float) * off);
    q->wbf.ensure(sizeof(unsigned short) * off);
    q->grad.ensure(sizeof(float) * off);
    q->m.ensure(sizeof(float) * off);
    q->v.ensure(sizeof(float) * off);
    MHIP_CHECK(hipMemsetAsync(q->m.ptr, 0, sizeof(float) * off, q->s));
    MHIP_CHECK(hipMemsetAsync(q->v.ptr, 0, sizeof(float) * off, q->s));
    MHIP_CHECK(hipStreamSynchronize(q->s));
    *out = q;
  });
}

void mhip_mlp_destroy(mhip_mlp* q) {
  if (!q) return;
  (void)hipStreamSynchronize(q->s);
  for (DevBuf* b : {&q->w32, &q->wbf, &q->grad, &q->m, &q->v, &q->acts, &q->xbf, &q->dbuf0, &q->dbuf1,
                    &q->tgt, &q->part, &q->loss})
    b->release();
  (void)hipStreamDestroy(q->s);
  delete q;
}

int mhip_mlp_set_params(mhip_mlp* q, const float* params) {
  return guard([&] {
    MHIP_CHECK(hipMemcpyAsync(q->w32.ptr, params, sizeof(float) * q->np, hipMemcpyHostToDevice, q->s));
    refresh_bf16(q);
    MHIP_CHECK(hipMemsetAsync(q->m.ptr, 0, sizeof(float) * q->np, q->s));
    MHIP_CHECK(hipMemsetAsync(q->v.ptr, 0, sizeof(float) * q->np, q->s));
    q->step = 0;
    MHIP_CHECK(hipStreamSynchronize(q->s));
  });
}

int mhip_mlp_get_params(mhip_mlp* q, float* params) {
  return guard([&] {
    MHIP_CHECK(hipMemcpyAsync(params, q->w32.ptr, sizeof(float) * q->np, hipMemcpyDeviceToHost, q->s));
    MHIP_CHECK(hipStreamSynchronize(q->s));
  });
}

int mhip_mlp_forward(mhip_mlp* q, int batch, const float* x, float* y) {
  return guard([&] {
    MHIP_REQUIRE(batch > 0, "mhip_mlp_forward: empty batch");
    mlp_ensure_batch(q, batch);
    MHIP_CHECK(hipMemcpyAsync(act_ptr(q, batch, 0), x, sizeof(float) * (size_t)batch * q->dims[0],
                              hipMemcpyHostToDevice, q->s));
    forward(q, batch);
    MHIP_CHECK(hipMemcpyAsync(y, act_ptr(q, batch, q->nl), sizeof(float) * (size_t)batch * q->dims[q->nl],
                              hipMemcpyDeviceToHost, q->s));
    MHIP_CHECK(hipStreamSynchronize(q->s));
  });
}

int mhip_mlp_gradients(mhip_mlp* q, int batch, const float* x, const float* t, float* grad,
                       float* loss) {
  return guard([&] {
    MHIP_REQUIRE(batch > 0, "mhip_mlp_gradients: empty batch");
    mlp_ensure_batch(q, batch);
    MHIP_CHECK(hipMemcpyAsync(act_ptr(q, batch, 0), x, sizeof(float) * (size_t)batch * q->dims[0],
                              hipMemcpyHostToDevice, q->s));
    MHIP_CHECK(hipMemcpyAsync(q->tgt.ptr, t, sizeof(float) * (size_t)batch * q->dims[q->nl],
                              hipMemcpyHostToDevice, q->s));
    forward(q, batch);
    backward(q, batch);
    MHIP_CHECK(hipMemcpyAsync(grad, q->grad.ptr, sizeof(float) * q->np, hipMemcpyDeviceToHost, q->s));
    MHIP_CHECK(hipMemcpyAsync(loss, q->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, q->s));
    MHIP_CHECK(hipStreamSynchronize(q->s));
  });
}

int mhip_mlp_train_step(mhip_mlp* q, int batch, const float* x, const float* t, float lr,
                        float* loss) {
  return guard([&] {
    MHIP_REQUIRE(batch > 0, "mhip_mlp_train_step: empty batch");
    mlp_ensure_batch(q, batch);
    MHIP_CHECK(hipMemcpyAsync(act_ptr(q, batch, 0), x, sizeof(float) * (size_t)batch * q->dims[0],
                              hipMemcpyHostToDevice, q->s));
    MHIP_CHECK(hipMemcpyAsync(q->tgt.ptr, t, sizeof(float) * (size_t)batch * q->dims[q->nl],
                              hipMemcpyHostToDevice, q->s));
    forward(q, batch);
    backward(q, batch);
    q->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)q->step), c2 = 1.f - std::pow(b2, (float)q->step);
    hipLaunchKernelGGL(k_adam, dim3(std::min(4096, cdiv(q->np, 256))), dim3(256), 0, q->s,
                       (long long)q->np, q->w32.as<float>(), q->grad.as<float>(), q->m.as<float>(),
                       q->v.as<float>(), lr, b1, b2, c1, c2, 1e-8f);
    MHIP_LAUNCH_CHECK();
    refresh_bf16(q);
    if (loss) {
      MHIP_CHECK(hipMemcpyAsync(loss, q->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, q->s));
    }
    MHIP_CHECK(hipStreamSynchronize(q->s));
  });
}

}  // extern "C"
