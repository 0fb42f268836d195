// Device primitives of the surrogate MLP (BF16 weights and activations, FP32 accumulation on
// the gfx950 matrix cores) shared by its training and inference kernels.
//
// Network of one ensemble member (rows = jump environments):
//   x   : one-hot of the species at the 60 environment slots, one channel per species that is
//         neither the base (solvent) species nor the vacancy; k = slot * C + channel, padded
//         with zeros to Kp (a multiple of 32). The base and vacancy channels of the full
//         one-hot input carry first-layer weights tied to zero: the channels of a slot sum to
//         one, so this loses nothing, and the vacancy channel is constant (the vacancy slot)
//         or out of distribution.
//   z1 = x W1 + b1, h1 = bf16(silu(z1))           W1 [Kp][128]
//   z2 = h1 W2 + b2, h2 = bf16(silu(z2))          W2 [128][128]
//   y  = h2 W3 + b3                               W3 [128][2]  -> (E_kra, dE), normalised
// Matrix products use v_mfma_f32_16x16x32_bf16 on 16-row tiles. Fragment layout (64 lanes):
//   A (16 x 32): lane l holds row l % 16, k = 8 (l / 16) .. +7
//   B (32 x 16): lane l holds column l % 16, k = 8 (l / 16) .. +7
//   C (16 x 16): lane l holds rows 4 (l / 16) + i (i < 4), column l % 16
// (A and B place k identically, so the product does not depend on the hardware's k order.)
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lmc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kHidden = 128;
constexpr int kMlpOut = 2;
constexpr int kEnvRow = 64;           // bytes per environment row (60 slots, padded)
constexpr int kEnvUsed = 60;
constexpr uint32_t kBf16One = 0x3F80u;

// Parameter layout of one member (FP32 master copy, gradients): W1 [Kp][128] | b1 [128] |
// W2 [128][128] | b2 [128] | W3 [128][2] | b3 [2], padded to a multiple of 4.
struct MlpLayout {
  int Kp, oW1, ob1, oW2, ob2, oW3, ob3, P;
};
__host__ __device__ inline MlpLayout MakeMlpLayout(int Kp) {
  MlpLayout L;
  L.Kp = Kp;
  L.oW1 = 0;
  L.ob1 = Kp * kHidden;
  L.oW2 = L.ob1 + kHidden;
  L.ob2 = L.oW2 + kHidden * kHidden;
  L.oW3 = L.ob2 + kHidden;
  L.ob3 = L.oW3 + kHidden * kMlpOut;
  L.P = (L.ob3 + kMlpOut + 3) & ~3;
  return L;
}

// Kernel view of an ensemble.
struct MlpDev {
  int M, C, Kp;
  MlpLayout lay;
  int chan[4];              // species code of each input channel
  const float *master;      // [M][P] FP32 parameters (biases are read from here)
  const uint16_t *w1t;      // [M][128][Kp] bf16 of W1, column-major (n, k)
  const uint16_t *w2t;      // [M][128][128] bf16 of W2 as (n, k)
  const uint16_t *w2k;      // [M][128][128] bf16 of W2 as (k, n)
  const uint16_t *w3;       // [M][128][2] bf16
  // the same W1 and W2 in the inference kernel's fragment order, one 16-byte fragment per lane
  // and 1 KB per (unit tile, k-step): w1f [M][8][Kp/32][64][8], w2f [M][8][4][64][8] (W2 in
  // the chained k order of the transposed first layer's accumulators)
  const uint16_t *w1f, *w2f;
};

// Round to nearest even (finite inputs).
__host__ __device__ inline uint32_t Bf16Bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}
__host__ __device__ inline float Bf16Float(uint32_t b) { return __builtin_bit_cast(float, b << 16); }
__host__ __device__ inline float Bf16Round(float f) { return Bf16Float(Bf16Bits(f)); }

__device__ inline float Sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

__device__ inline f32x4 Mfma(const u32x4 &a, const u32x4 &b, const f32x4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// 16 bytes (8 bf16) at p (16-byte aligned), global or LDS.
__device__ inline u32x4 Load8(const uint16_t *p) { return *reinterpret_cast<const u32x4 *>(p); }

// A fragment of the one-hot input: row environment e (kEnvRow bytes), k = k0 .. k0+7.
template <int C>
__device__ inline u32x4 OneHotFragment(const uint8_t *e, int k0, const int *chan) {
  u32x4 f;
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    uint32_t w = 0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int k = k0 + 2 * h + b;
      const int j = k / C, c = k - j * C;
      const bool on = j < kEnvUsed && e[j] == chan[c];
      w |= (on ? kBf16One : 0u) << (16 * b);
    }
    f[h] = w;
  }
  return f;
}

// Packs 8 bf16 from 8 floats already rounded to bf16.
__device__ inline u32x4 PackBf16(const float *v) {
  u32x4 f;
#pragma unroll
  for (int h = 0; h < 4; ++h) f[h] = Bf16Bits(v[2 * h]) | (Bf16Bits(v[2 * h + 1]) << 16);
  return f;
}

}  // namespace lmc
