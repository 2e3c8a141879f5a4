// Learned starting density for water clusters, on the device.
//
// Stages (one launch each, all on the caller's stream except the network):
//   k_gw_frames    per molecule: its frame (bisector, in-plane, normal) and
//                  internal geometry (two OH lengths, HOH cosine)
//   k_gw_atoms     per atom, over every atom of the other molecules: radial
//                  Gaussian densities of O and H neighbours (r <= 4 A,
//                  cosine cutoff) and, in the atom's molecule frame, the
//                  potential, field and field gradient of the other
//                  molecules' PM7 water charges (r <= 9 A, switched off over
//                  the last 1 A) plus unit-vector sums toward O and H
//                  neighbours in two distance shells
//   k_gw_pairs     per atom pair (A >= B): type one-hot, relative position and
//                  orientation, both atoms' descriptors; normalized and
//                  written as FP32 straight into the network's input
//   MLP forward    3 x 256 tanh layers, BF16 weights, FP32 accumulation
//   k_gw_assemble  per pair: de-normalize the predicted 4 x 4 block, rotate
//                  it from the molecule frames to the molecular axes and
//                  write it into the packed density
//   k_gw_norm      per molecule: rescale its diagonal to 8 valence electrons
// The SCF then starts from this density as iter starts from an OLDENS
// density (src/SCF/iter.F90:151-159).  Every packed slot is written by the
// one pair that owns it; slots of pairs beyond the neighbour list stay zero.
#include <cmath>
#include <vector>

#include "k_guess.h"

namespace mhip {
namespace {

constexpr double kRcut = 4.0;
constexpr double kEta = 3.0;
constexpr double kRfield = 9.0;
constexpr double kQO = -0.644520, kQH = 0.322260;
constexpr int kNmu = 6;           // radial Gaussians at 1.0, 1.6, ..., 4.0 A
constexpr int kNrad = 2 * kNmu;   // O then H neighbours
constexpr int kNenv = 9 + 12;     // potential, field (3), gradient (5), shell sums (2 x 2 x 3)
constexpr int kNatomf = kNrad + kNenv;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void k_gw_frames(int nw, const int* __restrict__ waters, const double* __restrict__ xyz,
                            double* __restrict__ R, double* __restrict__ geo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  const double* o = xyz + 3 * waters[3 * k];
  const double* h1 = xyz + 3 * waters[3 * k + 1];
  const double* h2 = xyz + 3 * waters[3 * k + 2];
  double v1[3], v2[3], e1[3], e2[3], d[3];
  for (int i = 0; i < 3; ++i) {
    v1[i] = h1[i] - o[i];
    v2[i] = h2[i] - o[i];
    e1[i] = v1[i] + v2[i];
    d[i] = h1[i] - h2[i];
  }
  const double nb = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
  for (int i = 0; i < 3; ++i) e1[i] /= nb;
  const double de = d[0] * e1[0] + d[1] * e1[1] + d[2] * e1[2];
  for (int i = 0; i < 3; ++i) d[i] -= de * e1[i];
  const double nd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int i = 0; i < 3; ++i) e2[i] = d[i] / nd;
  double* r = R + 9 * k;
  for (int i = 0; i < 3; ++i) {
    r[i] = e1[i];
    r[3 + i] = e2[i];
  }
  r[6] = e1[1] * e2[2] - e1[2] * e2[1];
  r[7] = e1[2] * e2[0] - e1[0] * e2[2];
  r[8] = e1[0] * e2[1] - e1[1] * e2[0];
  const double r1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
  const double r2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
  const double ca = (v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) / (r1 * r2);
  geo[3 * k] = r1 - 0.96;
  geo[3 * k + 1] = r2 - 0.96;
  geo[3 * k + 2] = ca - cos(104.5 * M_PI / 180.0);
}

__global__ __launch_bounds__(256) void k_gw_atoms(int numat, const int* __restrict__ nat,
                                                  const double* __restrict__ xyz, const int* __restrict__ mol,
                                                  const double* __restrict__ R, double* __restrict__ atomf) {
  const int a = blockIdx.x;
  __shared__ double red[4][kNatomf];
  double acc[kNatomf];
#pragma unroll
  for (int k = 0; k < kNatomf; ++k) acc[k] = 0.0;
  const int ma = mol[a];
  double Ra[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Ra[k] = R[9 * ma + k];
  const double xa = xyz[3 * a], ya = xyz[3 * a + 1], za = xyz[3 * a + 2];
  for (int b = threadIdx.x; b < numat; b += 256) {
    if (mol[b] == ma) continue;
    const double d0 = xyz[3 * b] - xa, d1 = xyz[3 * b + 1] - ya, d2 = xyz[3 * b + 2] - za;
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    if (r > kRfield) continue;
    const bool oxy = nat[b] == 8;
    if (r <= kRcut) {
      const double fc = 0.5 * (cos(M_PI * r / kRcut) + 1.0);
      const int c0 = oxy ? 0 : kNmu;
#pragma unroll
      for (int m = 0; m < kNmu; ++m) {
        const double mu = 1.0 + 0.6 * m;
        const double g = exp(-kEta * (r - mu) * (r - mu)) * fc;
        if (c0 == 0) acc[m] += g;
        else acc[kNmu + m] += g;
      }
    }
    const double e0 = d0 / r, e1 = d1 / r, e2 = d2 / r;
    const double u0 = Ra[0] * e0 + Ra[1] * e1 + Ra[2] * e2;
    const double u1 = Ra[3] * e0 + Ra[4] * e1 + Ra[5] * e2;
    const double u2 = Ra[6] * e0 + Ra[7] * e1 + Ra[8] * e2;
    const double sw = r < kRfield - 1.0 ? 1.0 : 0.5 * (1.0 + cos(M_PI * (r - (kRfield - 1.0))));
    const double qs = (oxy ? kQO : kQH) * sw;
    double* env = acc + kNrad;
    env[0] += qs / r;
    const double f = qs / (r * r);
    env[1] += f * u0;
    env[2] += f * u1;
    env[3] += f * u2;
    const double g = qs / (r * r * r);
    env[4] += g * (3.0 * u0 * u0 - 1.0);
    env[5] += g * (3.0 * u1 * u1 - 1.0);
    env[6] += g * 3.0 * u0 * u1;
    env[7] += g * 3.0 * u0 * u2;
    env[8] += g * 3.0 * u1 * u2;
    // shells [1.4, 2.6) and [2.6, 4.0) A, O then H
    const int si = (r >= 1.4 && r < 2.6) ? 0 : ((r >= 2.6 && r < 4.0) ? 1 : -1);
    if (si >= 0) {
      const int c0 = 9 + 3 * (2 * si + (oxy ? 0 : 1));
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (9 + 3 * q == c0) {
          env[9 + 3 * q] += u0;
          env[10 + 3 * q] += u1;
          env[11 + 3 * q] += u2;
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kNatomf; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < kNatomf) {
    const int k = threadIdx.x;
    atomf[(size_t)a * kNatomf + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

// From here on no multiply-add contraction: the descriptor normalization and
// the de-normalization of the network outputs round like the host's FP32
// arithmetic (a product, then a sum).
#pragma clang fp contract(off)

__device__ __forceinline__ int pair_type(bool same, bool intra, bool oa, bool ob) {
  if (same) return oa ? 0 : 1;
  if (intra) return (oa || ob) ? 2 : 3;
  if (oa && ob) return 4;
  if (oa) return 5;
  if (ob) return 6;
  return 7;
}

__global__ __launch_bounds__(128) void k_gw_pairs(
    int npairs, const int* __restrict__ pa, const int* __restrict__ pb, const int* __restrict__ nat,
    const int* __restrict__ mol, const double* __restrict__ side, const double* __restrict__ xyz,
    const double* __restrict__ R, const double* __restrict__ geo, const double* __restrict__ atomf,
    const float* __restrict__ xm, const float* __restrict__ xs, float* __restrict__ X, int* __restrict__ typ) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npairs) return;
  const int a = pa[m], b = pb[m];
  const int ma = mol[a], mb = mol[b];
  const bool same = a == b, intra = !same && ma == mb;
  const int t = pair_type(same, intra, nat[a] == 8, nat[b] == 8);
  typ[m] = t;
  float* x = X + (size_t)m * kGuessNin;
  auto put = [&](int k, double v) {
    const float v32 = (float)v;
    x[k] = (v32 - xm[k]) / xs[k];
  };
  for (int k = 0; k < 8; ++k) put(k, k == t ? 1.0 : 0.0);
  const double* Ra = R + 9 * ma;
  const double* Rb = R + 9 * mb;
  const double w0 = xyz[3 * b] - xyz[3 * a], w1 = xyz[3 * b + 1] - xyz[3 * a + 1],
               w2 = xyz[3 * b + 2] - xyz[3 * a + 2];
  double d[3];
  for (int i = 0; i < 3; ++i) d[i] = Ra[3 * i] * w0 + Ra[3 * i + 1] * w1 + Ra[3 * i + 2] * w2;
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int i = 0; i < 3; ++i) put(8 + i, d[i] / 4.0);
  put(11, r / 4.0);
  put(12, r > 1e-6 ? 1.0 / r : 0.0);
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k)
      put(13 + 3 * i + k, Ra[3 * i] * Rb[3 * k] + Ra[3 * i + 1] * Rb[3 * k + 1] + Ra[3 * i + 2] * Rb[3 * k + 2]);
  const double* fa = atomf + (size_t)a * kNatomf;
  const double* fb = atomf + (size_t)b * kNatomf;
  for (int k = 0; k < kNrad; ++k) {
    put(22 + k, fa[k]);
    put(34 + k, fb[k]);
  }
  for (int k = 0; k < 3; ++k) {
    put(46 + k, geo[3 * ma + k]);
    put(49 + k, geo[3 * mb + k]);
  }
  put(52, side[a]);
  put(53, side[b]);
  for (int k = 0; k < kNenv; ++k) {
    put(54 + k, fa[kNrad + k]);
    put(75 + k, fb[kNrad + k]);
  }
}

// orbital rotation of one atom: U = 1 (+) R^T for oxygen (s, p), 1 for hydrogen
__device__ __forceinline__ double uel(const double* Rm, bool oxy, int i, int x) {
  if (i == 0 || x == 0) return (i == x) ? 1.0 : 0.0;
  if (!oxy) return (i == x) ? 1.0 : 0.0;
  return Rm[3 * (x - 1) + (i - 1)];  // (R^T)[i-1][x-1]
}

__global__ __launch_bounds__(128) void k_gw_assemble(
    int npairs, const int* __restrict__ pa, const int* __restrict__ pb, const int* __restrict__ nat,
    const int* __restrict__ mol, const int* __restrict__ nf0, const int* __restrict__ no,
    const double* __restrict__ R, const int* __restrict__ typ, const float* __restrict__ Y,
    const float* __restrict__ ym, const float* __restrict__ ys, double* __restrict__ p) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npairs) return;
  const int a = pa[m], b = pb[m], t = typ[m];
  double L[4][4];
  const float* y = Y + (size_t)m * kGuessNout;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * i + j;
      L[i][j] = (double)(y[k] * ys[t * kGuessNout + k] + ym[t * kGuessNout + k]);
    }
  const bool same = a == b;
  if (same)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < i; ++j) L[i][j] = L[j][i] = 0.5 * (L[i][j] + L[j][i]);
  const bool oa = nat[a] == 8, ob = nat[b] == 8;
  const double* Ra = R + 9 * mol[a];
  const double* Rb = R + 9 * mol[b];
  const int na = no[a], nb = no[b], ra0 = nf0[a], cb0 = nf0[b];
  for (int i = 0; i < na; ++i) {
    // T = U_a L  (row i)
    double T[4];
    for (int y2 = 0; y2 < 4; ++y2) {
      double s = 0.0;
      for (int x = 0; x < 4; ++x) s += uel(Ra, oa, i, x) * L[x][y2];
      T[y2] = s;
    }
    for (int j = 0; j < nb; ++j) {
      if (same && j > i) break;
      double s = 0.0;
      for (int y2 = 0; y2 < 4; ++y2) s += T[y2] * uel(Rb, ob, j, y2);
      const long long r = ra0 + i, c = cb0 + j;
      p[r * (r + 1) / 2 + c] = s;
    }
  }
}

__global__ void k_gw_norm(int nw, const int* __restrict__ waters, const int* __restrict__ nf0,
                          const int* __restrict__ no, double* __restrict__ p) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  double tot = 0.0;
  for (int q = 0; q < 3; ++q) {
    const int at = waters[3 * k + q];
    for (int i = 0; i < no[at]; ++i) {
      const long long o = nf0[at] + i;
      tot += p[o * (o + 1) / 2 + o];
    }
  }
  const double sc = tot > 1e-6 ? 8.0 / tot : 1.0;
  for (int q = 0; q < 3; ++q) {
    const int at = waters[3 * k + q];
    for (int i = 0; i < no[at]; ++i) {
      const long long o = nf0[at] + i;
      p[o * (o + 1) / 2 + o] *= sc;
    }
  }
}

void put(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
  b.ensure(bytes ? bytes : 8);
  if (bytes) MHIP_CHECK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s));
}

inline int blocks(long long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

void guess_water(hipStream_t s, GuessWork& g, mhip_mlp* q, const GuessHost& in, double* p, long long mpack,
                 float* x_out, float* y_out) {
  const int n = in.numat, nw = in.nw, np = in.npairs;
  std::vector<double> side(n, 0.0);
  for (int k = 0; k < nw; ++k) {
    side[in.waters[3 * k + 1]] = 1.0;
    side[in.waters[3 * k + 2]] = -1.0;
  }
  put(g.xyz, in.xyz, sizeof(double) * 3 * n, s);
  put(g.nat, in.nat, sizeof(int) * n, s);
  put(g.mol, in.mol, sizeof(int) * n, s);
  put(g.side, side.data(), sizeof(double) * n, s);
  put(g.waters, in.waters, sizeof(int) * 3 * nw, s);
  put(g.pa, in.pa, sizeof(int) * np, s);
  put(g.pb, in.pb, sizeof(int) * np, s);
  put(g.nf0, in.nf0, sizeof(int) * n, s);
  put(g.no, in.no, sizeof(int) * n, s);
  std::vector<float> norm;
  norm.insert(norm.end(), in.x_mean, in.x_mean + kGuessNin);
  norm.insert(norm.end(), in.x_std, in.x_std + kGuessNin);
  norm.insert(norm.end(), in.y_mean, in.y_mean + kGuessTypes * kGuessNout);
  norm.insert(norm.end(), in.y_std, in.y_std + kGuessTypes * kGuessNout);
  put(g.norm, norm.data(), sizeof(float) * norm.size(), s);
  const float* xm = g.norm.as<float>();
  const float* xs = xm + kGuessNin;
  const float* ymn = xs + kGuessNin;
  const float* ysd = ymn + kGuessTypes * kGuessNout;
  g.R.ensure(sizeof(double) * 9 * nw);
  g.geo.ensure(sizeof(double) * 3 * nw);
  g.atomf.ensure(sizeof(double) * kNatomf * n);
  g.typ.ensure(sizeof(int) * np);

  hipLaunchKernelGGL(k_gw_frames, dim3(blocks(nw, 64)), dim3(64), 0, s, nw, g.waters.as<int>(),
                     g.xyz.as<double>(), g.R.as<double>(), g.geo.as<double>());
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gw_atoms, dim3(n), dim3(256), 0, s, n, g.nat.as<int>(), g.xyz.as<double>(),
                     g.mol.as<int>(), g.R.as<double>(), g.atomf.as<double>());
  MHIP_LAUNCH_CHECK();
  float* X = mlp_input(q, np);
  hipLaunchKernelGGL(k_gw_pairs, dim3(blocks(np, 128)), dim3(128), 0, s, np, g.pa.as<int>(), g.pb.as<int>(),
                     g.nat.as<int>(), g.mol.as<int>(), g.side.as<double>(), g.xyz.as<double>(), g.R.as<double>(),
                     g.geo.as<double>(), g.atomf.as<double>(), xm, xs, X, g.typ.as<int>());
  MHIP_LAUNCH_CHECK();
  // the network runs on its own stream: hand the inputs over, take the outputs back
  MHIP_CHECK(hipStreamSynchronize(s));
  const float* Y = mlp_forward_resident(q, np);
  fill(s, p, mpack, 0.0);
  hipLaunchKernelGGL(k_gw_assemble, dim3(blocks(np, 128)), dim3(128), 0, s, np, g.pa.as<int>(), g.pb.as<int>(),
                     g.nat.as<int>(), g.mol.as<int>(), g.nf0.as<int>(), g.no.as<int>(), g.R.as<double>(),
                     g.typ.as<int>(), Y, ymn, ysd, p);
  MHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gw_norm, dim3(blocks(nw, 64)), dim3(64), 0, s, nw, g.waters.as<int>(), g.nf0.as<int>(),
                     g.no.as<int>(), p);
  MHIP_LAUNCH_CHECK();
  if (x_out)
    MHIP_CHECK(hipMemcpyAsync(x_out, X, sizeof(float) * (size_t)np * kGuessNin, hipMemcpyDeviceToHost, s));
  if (y_out)
    MHIP_CHECK(hipMemcpyAsync(y_out, Y, sizeof(float) * (size_t)np * kGuessNout, hipMemcpyDeviceToHost, s));
  MHIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace mhip
