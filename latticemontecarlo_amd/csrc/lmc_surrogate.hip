// gfx950 kernels of the surrogate ensemble: training (fused forward + backward per 64-row tile,
// deterministic per-tile gradient partials, fixed-order reduction + Adam) and the forward pass.
// The network and fragment layouts are described in lmc_mlp.h.
#include <hip/hip_runtime.h>

#include "lmc_mlp.h"
#include "lmc_surrogate.h"

namespace lmc {
namespace {

constexpr int kRows = 64;  // rows per training workgroup (one 16-row tile per wave)

template <int C>
struct KDims {
  static constexpr int Kp = (kEnvUsed * C + 31) / 32 * 32;
  static constexpr int KS = Kp / 32;
};

struct TrainLds {
  uint8_t env[kRows][kEnvRow];
  uint16_t h1[kRows][kHidden];
  uint16_t h1t[kHidden][kRows];
  uint16_t dz2[kRows][kHidden];
  uint16_t dz2t[kHidden][kRows];
  uint16_t dz1t[kHidden][kRows];
  float tg[kRows][kMlpOut];
  float rw3[4][kHidden][kMlpOut];
  float rb1[4][kHidden], rb2[4][kHidden];
  float rb3[4][kMlpOut], rloss[4];
};

// Loads the tile's environment rows (0xFF rows past the end: no channel matches) and targets.
__device__ void LoadRows(TrainLds &S, const uint8_t *env, const int32_t *order, int row0, int n,
                         const float *tgt) {
  const int tid = threadIdx.x, r = tid >> 2, part = tid & 3;
  const int p = blockIdx.x * kRows + r;
  u32x4 v = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
  int s = 0;
  if (p < n) {
    s = order ? order[row0 + p] : row0 + p;
    v = *reinterpret_cast<const u32x4 *>(env + static_cast<size_t>(s) * kEnvRow + 16 * part);
  }
  *reinterpret_cast<u32x4 *>(&S.env[r][16 * part]) = v;
  if (tgt && part < kMlpOut) S.tg[r][part] = p < n ? tgt[static_cast<size_t>(s) * kMlpOut + part] : 0.f;
}

// Forward pass of the wave's 16-row tile: z1, z2 (C layout), h2 (bf16 values as float), y.
// Writes h1 (row-major and transposed) to LDS.
template <int C>
__device__ void TileForward(const MlpDev &w, int m, TrainLds &S, float (&z1)[8][4],
                            float (&z2)[8][4], float (&h2)[8][4], float (&y)[4][kMlpOut]) {
  constexpr int Kp = KDims<C>::Kp, KS = KDims<C>::KS;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, col = l & 15, grp = l >> 4;
  const int r0 = 16 * wv;
  const float *par = w.master + static_cast<size_t>(m) * w.lay.P;
  const uint16_t *W1t = w.w1t + static_cast<size_t>(m) * kHidden * Kp;
  const uint16_t *W2t = w.w2t + static_cast<size_t>(m) * kHidden * kHidden;
  const uint16_t *W3 = w.w3 + static_cast<size_t>(m) * kHidden * kMlpOut;
  int chan[4] = {w.chan[0], w.chan[1], w.chan[2], w.chan[3]};
  u32x4 a[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) a[s] = OneHotFragment<C>(S.env[r0 + col], 32 * s + 8 * grp, chan);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const uint16_t *bp = W1t + static_cast<size_t>(16 * t + col) * Kp + 8 * grp;
#pragma unroll
    for (int s = 0; s < KS; ++s) acc = Mfma(a[s], Load8(bp + 32 * s), acc);
    const float bias = par[w.lay.ob1 + 16 * t + col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      z1[t][i] = acc[i] + bias;
      const uint32_t hb = Bf16Bits(z1[t][i] * Sigmoid(z1[t][i]));
      const int r = r0 + 4 * grp + i, n = 16 * t + col;
      S.h1[r][n] = static_cast<uint16_t>(hb);
      S.h1t[n][r] = static_cast<uint16_t>(hb);
    }
  }
  __syncthreads();
  u32x4 b[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) b[s] = Load8(&S.h1[r0 + col][32 * s + 8 * grp]);
  float p[4][kMlpOut] = {};
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const uint16_t *bp = W2t + static_cast<size_t>(16 * t + col) * kHidden + 8 * grp;
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = Mfma(b[s], Load8(bp + 32 * s), acc);
    const float bias = par[w.lay.ob2 + 16 * t + col];
    const float w30 = Bf16Float(W3[(16 * t + col) * kMlpOut]);
    const float w31 = Bf16Float(W3[(16 * t + col) * kMlpOut + 1]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      z2[t][i] = acc[i] + bias;
      h2[t][i] = Bf16Round(z2[t][i] * Sigmoid(z2[t][i]));
      p[i][0] += h2[t][i] * w30;
      p[i][1] += h2[t][i] * w31;
    }
  }
  // sum over the 16 columns held by the lanes of this row group (identical in every lane)
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) {
      float v = p[i][o];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      y[i][o] = v + par[w.lay.ob3 + o];
    }
}

// Sum over the four row groups (lanes l, l^16, l^32, l^48); identical in all four.
__device__ inline float GroupSum(float v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

template <int C>
__global__ __launch_bounds__(256) void MlpTrainKernel(MlpDev w, const uint8_t *env,
                                                      const int32_t *order, int row0, int n,
                                                      const float *tgt, float inv_n,
                                                      float *grad_part, float *loss_part) {
  constexpr int Kp = KDims<C>::Kp;
  __shared__ TrainLds S;
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6, col = l & 15, grp = l >> 4;
  const int r0 = 16 * wv, m = blockIdx.y;
  const MlpLayout &Ly = w.lay;
  LoadRows(S, env, order, row0, n, tgt);
  __syncthreads();
  float z1[8][4], z2[8][4], h2[8][4], y[4][kMlpOut];
  TileForward<C>(w, m, S, z1, z2, h2, y);
  const uint16_t *W2k = w.w2k + static_cast<size_t>(m) * kHidden * kHidden;
  const uint16_t *W3 = w.w3 + static_cast<size_t>(m) * kHidden * kMlpOut;

  // loss and dL/dy (L = sum over rows and outputs of (y - t)^2 / (2 n))
  float dy[4][kMlpOut], lsum = 0.f, db3[kMlpOut] = {0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + 4 * grp + i;
    const bool valid = blockIdx.x * kRows + r < n;
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) {
      const float d = y[i][o] - S.tg[r][o];
      dy[i][o] = valid ? d * inv_n : 0.f;
      lsum += valid ? d * d : 0.f;
      db3[o] += dy[i][o];
    }
  }
  lsum = GroupSum(lsum);
  db3[0] = GroupSum(db3[0]);
  db3[1] = GroupSum(db3[1]);
  if (l == 0) {
    S.rloss[wv] = lsum;
    S.rb3[wv][0] = db3[0];
    S.rb3[wv][1] = db3[1];
  }
  // dW3, db2 and dz2 = (dy W3^T) * silu'(z2)
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int nn = 16 * t + col;
    const float w30 = Bf16Float(W3[nn * kMlpOut]), w31 = Bf16Float(W3[nn * kMlpOut + 1]);
    float g0 = 0.f, g1 = 0.f, gb = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      g0 += h2[t][i] * dy[i][0];
      g1 += h2[t][i] * dy[i][1];
      const float dh = dy[i][0] * w30 + dy[i][1] * w31;
      const float sg = Sigmoid(z2[t][i]);
      const float dz = dh * sg * (1.f + z2[t][i] * (1.f - sg));
      gb += dz;
      const uint32_t bb = Bf16Bits(dz);
      const int r = r0 + 4 * grp + i;
      S.dz2[r][nn] = static_cast<uint16_t>(bb);
      S.dz2t[nn][r] = static_cast<uint16_t>(bb);
    }
    g0 = GroupSum(g0);
    g1 = GroupSum(g1);
    gb = GroupSum(gb);
    if (grp == 0) {
      S.rw3[wv][nn][0] = g0;
      S.rw3[wv][nn][1] = g1;
      S.rb2[wv][nn] = gb;
    }
  }
  __syncthreads();
  // dh1 = dz2 W2^T, dz1 = dh1 * silu'(z1)
  {
    u32x4 a[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) a[s] = Load8(&S.dz2[r0 + col][32 * s + 8 * grp]);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const uint16_t *bp = W2k + static_cast<size_t>(16 * t + col) * kHidden + 8 * grp;
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = Mfma(a[s], Load8(bp + 32 * s), acc);
      const int nn = 16 * t + col;
      float gb = 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float sg = Sigmoid(z1[t][i]);
        const float dz = acc[i] * sg * (1.f + z1[t][i] * (1.f - sg));
        gb += dz;
        S.dz1t[nn][r0 + 4 * grp + i] = static_cast<uint16_t>(Bf16Bits(dz));
      }
      gb = GroupSum(gb);
      if (grp == 0) S.rb1[wv][nn] = gb;
    }
  }
  __syncthreads();
  float *g = grad_part + (static_cast<size_t>(blockIdx.x) * w.M + m) * Ly.P;
  // dW2 = h1^T dz2 (rows k of W2 from h1^T, columns n from dz2^T; K = the tile's 64 rows)
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int kt = 2 * wv + u;
    u32x4 a[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) a[s] = Load8(&S.h1t[16 * kt + col][32 * s + 8 * grp]);
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 2; ++s) acc = Mfma(a[s], Load8(&S.dz2t[16 * nt + col][32 * s + 8 * grp]), acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) g[Ly.oW2 + (16 * kt + 4 * grp + i) * kHidden + 16 * nt + col] = acc[i];
    }
  }
  // dW1 = x^T dz1 (rows k of W1 rebuilt from the environments)
  int chan[4] = {w.chan[0], w.chan[1], w.chan[2], w.chan[3]};
  for (int kt = wv; kt < Kp / 16; kt += 4) {
    const int k = 16 * kt + col;
    const int j = k / C, c = k - j * C;
    u32x4 a[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const int r = 32 * s + 8 * grp + 2 * h + b;
          const bool on = j < kEnvUsed && S.env[r][j] == chan[c];
          v |= (on ? kBf16One : 0u) << (16 * b);
        }
        a[s][h] = v;
      }
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 2; ++s) acc = Mfma(a[s], Load8(&S.dz1t[16 * nt + col][32 * s + 8 * grp]), acc);
#pragma unroll
      for (int i = 0; i < 4; ++i) g[Ly.oW1 + (16 * kt + 4 * grp + i) * kHidden + 16 * nt + col] = acc[i];
    }
  }
  // biases and W3 in a fixed order over the four waves
  if (tid < kHidden) {
    g[Ly.ob1 + tid] = ((S.rb1[0][tid] + S.rb1[1][tid]) + S.rb1[2][tid]) + S.rb1[3][tid];
    g[Ly.ob2 + tid] = ((S.rb2[0][tid] + S.rb2[1][tid]) + S.rb2[2][tid]) + S.rb2[3][tid];
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o)
      g[Ly.oW3 + tid * kMlpOut + o] =
          ((S.rw3[0][tid][o] + S.rw3[1][tid][o]) + S.rw3[2][tid][o]) + S.rw3[3][tid][o];
  }
  if (tid < Ly.P - Ly.ob3) {
    float v = 0.f;
    if (tid < kMlpOut) v = ((S.rb3[0][tid] + S.rb3[1][tid]) + S.rb3[2][tid]) + S.rb3[3][tid];
    g[Ly.ob3 + tid] = v;
  }
  if (tid == 0)
    loss_part[static_cast<size_t>(blockIdx.x) * w.M + m] =
        ((S.rloss[0] + S.rloss[1]) + S.rloss[2]) + S.rloss[3];
}

template <int C>
__global__ __launch_bounds__(256) void MlpForwardKernel(MlpDev w, const uint8_t *env, int n,
                                                        float *out) {
  __shared__ TrainLds S;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, col = l & 15, grp = l >> 4;
  const int m = blockIdx.y;
  LoadRows(S, env, nullptr, 0, n, nullptr);
  __syncthreads();
  float z1[8][4], z2[8][4], h2[8][4], y[4][kMlpOut];
  TileForward<C>(w, m, S, z1, z2, h2, y);
  if (col == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = blockIdx.x * kRows + 16 * wv + 4 * grp + i;
      if (p < n)
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) out[(static_cast<size_t>(p) * w.M + m) * kMlpOut + o] = y[i][o];
    }
}

// Sums the per-tile gradient partials in tile order; optionally applies Adam to the FP32 master
// parameters; refreshes the bf16 copies from the master parameters.
__global__ void MlpReduceAdamKernel(MlpMut w, const float *grad_part, int nparts, float *grad_out,
                                    int update, float lr, float beta1, float beta2, float eps,
                                    float corr1, float corr2) {
  const int M = w.M, P = w.lay.P;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * P) return;
  const int m = idx / P, p = idx - m * P;
  float v = w.master[idx];
  if (grad_part) {
    float g = 0.f;
    for (int b = 0; b < nparts; ++b) g += grad_part[(static_cast<size_t>(b) * M + m) * P + p];
    if (grad_out) grad_out[idx] = g;
    if (update) {
      const float m1 = beta1 * w.adam_m[idx] + (1.f - beta1) * g;
      const float m2 = beta2 * w.adam_v[idx] + (1.f - beta2) * g * g;
      w.adam_m[idx] = m1;
      w.adam_v[idx] = m2;
      v -= lr * (m1 * corr1) / (sqrtf(m2 * corr2) + eps);
      w.master[idx] = v;
    }
  }
  const MlpLayout &L = w.lay;
  const uint16_t b = static_cast<uint16_t>(Bf16Bits(v));
  if (p < L.ob1) {
    const int k = p / kHidden, nn = p - k * kHidden;
    w.w1t[(static_cast<size_t>(m) * kHidden + nn) * L.Kp + k] = b;
    // fragment order: unit tile nn / 16, k-step k / 32, lane (k % 32 / 8) * 16 + nn % 16
    const int ks = k >> 5, kw = k & 31, lane = (kw >> 3) * 16 + (nn & 15);
    w.w1f[((static_cast<size_t>(m) * 8 + (nn >> 4)) * (L.Kp >> 5) + ks) * 512 + lane * 8 + (kw & 7)] = b;
  } else if (p >= L.oW2 && p < L.ob2) {
    const int q = p - L.oW2, k = q / kHidden, nn = q - k * kHidden;
    w.w2t[(static_cast<size_t>(m) * kHidden + nn) * kHidden + k] = b;
    w.w2k[(static_cast<size_t>(m) * kHidden + k) * kHidden + nn] = b;
    // chained order: input k = 32j + 4grp + e (e < 4) or 32j + 16 + 4grp + (e - 4)
    const int j = k >> 5, kw = k & 31, hi = kw >> 4, grp = (kw & 15) >> 2;
    const int e = 4 * hi + (kw & 3), lane = grp * 16 + (nn & 15);
    w.w2f[((static_cast<size_t>(m) * 8 + (nn >> 4)) * 4 + j) * 512 + lane * 8 + e] = b;
  } else if (p >= L.oW3 && p < L.ob3) {
    w.w3[static_cast<size_t>(m) * kHidden * kMlpOut + (p - L.oW3)] = b;
  }
}

}  // namespace

hipError_t LaunchMlpTrain(const MlpDev &w, const uint8_t *env, const int32_t *order, int row0,
                          int n, const float *tgt, float *grad_part, float *loss_part,
                          hipStream_t s) {
  const dim3 grid((n + kRows - 1) / kRows, w.M);
  const float inv_n = 1.0f / static_cast<float>(n);
  if (w.C == 2)
    hipLaunchKernelGGL(MlpTrainKernel<2>, grid, dim3(256), 0, s, w, env, order, row0, n, tgt,
                       inv_n, grad_part, loss_part);
  else if (w.C == 3)
    hipLaunchKernelGGL(MlpTrainKernel<3>, grid, dim3(256), 0, s, w, env, order, row0, n, tgt,
                       inv_n, grad_part, loss_part);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t LaunchMlpForward(const MlpDev &w, const uint8_t *env, int n, float *out, hipStream_t s) {
  const dim3 grid((n + kRows - 1) / kRows, w.M);
  if (w.C == 2)
    hipLaunchKernelGGL(MlpForwardKernel<2>, grid, dim3(256), 0, s, w, env, n, out);
  else if (w.C == 3)
    hipLaunchKernelGGL(MlpForwardKernel<3>, grid, dim3(256), 0, s, w, env, n, out);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t LaunchMlpReduceAdam(const MlpMut &w, const float *grad_part, int nparts,
                               float *grad_out, int update, const AdamStep &a, hipStream_t s) {
  const int total = w.M * w.lay.P;
  hipLaunchKernelGGL(MlpReduceAdamKernel, dim3((total + 255) / 256), dim3(256), 0, s, w, grad_part,
                     nparts, grad_out, update, a.lr, a.beta1, a.beta2, a.eps, a.corr1, a.corr2);
  return hipGetLastError();
}

int MlpTileRows() { return kRows; }

}  // namespace lmc
