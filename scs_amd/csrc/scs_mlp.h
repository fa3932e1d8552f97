// Branching-score MLP (6 integer features -> 64 -> 64 -> 1), device side.
//
// The learned ordering replaces the connection score of selectUstar
// (src/helpers.cc:1104-1168) when a query runs with a model.  It only
// changes which vertex is branched on first; every rule and bound is the
// same, so the optimum is unchanged.
//
// Inference runs on the MFMA units: v_mfma_f32_16x16x16_bf16 with BF16
// weights/activations and FP32 accumulation, 16 candidates per tile.
// Operand layout of one 16x16x16 MFMA (lane l of 64, kb = l / 16):
//   A[m][k]: lane holds m = l % 16, k = 4*kb + i (i = 0..3)
//   B[k][n]: lane holds n = l % 16, k = 4*kb + i
//   D[m][n]: lane holds n = l % 16, m = 4*kb + i
#pragma once
#include "scs_internal.h"

#define MLP_IN 6
#define MLP_INP 16  // input padded to one MFMA K-step
#define MLP_H 64

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// Packed inference model in device memory.
struct MlpDev {
  u16 w1[MLP_INP][MLP_H];  // bf16, rows >= MLP_IN are zero
  u16 w2[MLP_H][MLP_H];    // bf16
  float b1[MLP_H];
  float b2[MLP_H];
  float w3[MLP_H];
  float b3;
  float mean[8];
  float scale[8];
  float pad[7];
};

__device__ __forceinline__ u16 f2bf(float f) {
  u32 x = __float_as_uint(f);
  x += 0x7fffu + ((x >> 16) & 1u);  // round to nearest even
  return (u16)(x >> 16);
}
__device__ __forceinline__ float bf2f(u16 b) { return __uint_as_float(((u32)b) << 16); }

// Raw features of a candidate v at the current node: four degrees/
// distances, the rule-1 slack and the plain connection score (in units).
struct MlpFeat {
  u32 dC, dR, dist, core;
  int slack;
  float cscore;
};

// Scores the 16 candidates of one tile.  feat(row) gives the raw features
// of row (rows >= valid are ignored).  h1s: 16 x 64 bf16 LDS scratch.
// Returns the score of row (l % 16) on every lane (lanes with the same
// l % 16 hold the same value).
template <typename FeatFn>
__device__ float mlp_tile(const MlpDev* M, FeatFn feat, u16* h1s) {
  const u32 l = threadIdx.x;
  const u32 row = l & 15u;
  const u32 kb = l >> 4;
  // layer 1 input: 16 rows x 16 k (features 0..5, rest zero)
  MlpFeat f = feat(row);
  float xin[6];
  xin[0] = (float)f.dC;
  xin[1] = (float)f.dR;
  xin[2] = (float)f.dist;
  xin[3] = (float)f.core;
  xin[4] = (float)f.slack;
  xin[5] = f.cscore;
  bf16x4 a1;
  for (int i = 0; i < 4; i++) {
    const u32 k = 4 * kb + i;
    float x = 0.f;
    if (k < MLP_IN) x = (xin[k] - M->mean[k]) * M->scale[k];
    a1[i] = (short)f2bf(x);
  }
#pragma unroll 1
  for (u32 j = 0; j < 4; j++) {
    bf16x4 b;
    for (int i = 0; i < 4; i++) b[i] = (short)M->w1[4 * kb + i][16 * j + row];
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b, d, 0, 0, 0);
    const u32 col = 16 * j + row;
    const float bias = M->b1[col];
    for (int i = 0; i < 4; i++) {
      float z = d[i] + bias;
      h1s[(4 * kb + i) * MLP_H + col] = f2bf(z > 0.f ? z : 0.f);
    }
  }
  __syncthreads();
  // layer 2 (K = 64 in four steps) fused with the 64 -> 1 output layer
  float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (u32 j = 0; j < 4; j++) {
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (u32 ks = 0; ks < 4; ks++) {
      bf16x4 a, b;
      for (int i = 0; i < 4; i++) {
        a[i] = (short)h1s[row * MLP_H + 16 * ks + 4 * kb + i];
        b[i] = (short)M->w2[16 * ks + 4 * kb + i][16 * j + row];
      }
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
    }
    const u32 col = 16 * j + row;
    const float bias = M->b2[col], w = M->w3[col];
    for (int i = 0; i < 4; i++) {
      float z = d[i] + bias;
      part[i] += (z > 0.f ? z : 0.f) * w;
    }
  }
  // part[i] is row (4*kb + i)'s partial sum over this lane's 4 columns;
  // sum over the 16 lanes of the same kb
  for (int i = 0; i < 4; i++)
    for (int o = 8; o > 0; o >>= 1) part[i] += __shfl_xor(part[i], o, 16);
  // lane l needs row l % 16 = 4*kb' + i'  -> read from a lane of group kb'
  // every lane of group kb now holds the sums of rows 4kb..4kb+3; lane l
  // needs row l % 16, held by group (l % 16) / 4 at index (l % 16) % 4
  const u32 src_kb = row >> 2, src_i = row & 3u;
  const float v0 = __shfl(part[0], 16 * src_kb), v1 = __shfl(part[1], 16 * src_kb);
  const float v2 = __shfl(part[2], 16 * src_kb), v3 = __shfl(part[3], 16 * src_kb);
  const float s = src_i == 0 ? v0 : src_i == 1 ? v1 : src_i == 2 ? v2 : v3;
  __syncthreads();
  return s + M->b3;
}
