// Learned eddy-viscosity closure: 7 -> 32 -> 32 -> 1 MLP (ReLU hidden layers,
// Softplus(beta=10) output), the 3-D counterpart of the reference's
// EddyViscosityMLP (src/tensorlbm/ai/model.py:40-92).
//
// Inference inside the collision runs on the matrix cores: per wave (64 cells,
// one per lane) the two hidden layers are six v_mfma_f32_32x32x16_bf16
// (bf16 operands, fp32 accumulation).  Layout (wave64, 32x32x16):
//   A[i][k] : lane i + 32*(k/8), element k%8
//   B[k][j] : lane j + 32*(k/8), element k%8
//   C[i][j] : lane j + 32*((i/4)%2), element (i%4) + 4*(i/8)
// Hidden units are rows (i) and cells are columns (j), so every lane feeds its
// own cell's features as B operand (layer 1 is split in two 32-cell blocks with
// the weights placed at K 0..7 or 8..15), the layer-1 C tile is reused as the
// layer-2 B operand in registers (W2's columns are permuted at packing time to
// match), and the output layer is 16 local FMAs plus one cross-half shuffle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tlbm {
namespace mlp {

constexpr int NIN = 7;
constexpr int NH = 32;
// fp32 parameter vector: W1 (32x7 row-major) b1 W2 (32x32) b2 w3 b3 | mean[7] inv_std[7]
constexpr int OW1 = 0;
constexpr int OB1 = OW1 + NH * NIN;  // 224
constexpr int OW2 = OB1 + NH;        // 256
constexpr int OB2 = OW2 + NH * NH;   // 1280
constexpr int OW3 = OB2 + NH;        // 1312
constexpr int OB3 = OW3 + NH;        // 1344
constexpr int NPARAM = OB3 + 1;      // 1345 trainable
constexpr int OMEAN = NPARAM;        // 1345
constexpr int OISTD = OMEAN + NIN;   // 1352
constexpr int NTOTAL = OISTD + NIN;  // 1359
constexpr float BETA = 10.0f;

// Matrix-core operand blob, built once per weight update by tlbm_mlp_pack.
struct Packed {
  uint16_t a1lo[64][8];   // layer-1 A for the cell block of lanes 0..31 (K 0..7)
  uint16_t a1hi[64][8];   // layer-1 A for the cell block of lanes 32..63 (K 8..15)
  uint16_t a2[2][64][8];  // layer-2 A, K-steps 0 and 1, columns in C-layout order
  float b1c[2][16];       // layer-1 bias, C layout per lane half
  float b2c[2][16];       // layer-2 bias
  float w3c[2][16];       // output weights
  float b3;
  float mean[7];
  float istd[7];
  float pad;
};
static_assert(sizeof(Packed) == 4544, "packed MLP blob layout");

// hidden unit held by C-tile element r of a lane in half h (= lane / 32)
__host__ __device__ constexpr int c_row(int h, int r) { return 4 * h + (r % 4) + 8 * (r / 4); }
// W2 column used by layer-2 K-step s, lane half h, element e
__host__ __device__ constexpr int w2_col(int s, int h, int e) { return 4 * h + (e % 4) + 16 * s + 8 * (e / 4); }

__device__ __forceinline__ uint16_t bf16_bits(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 as_bf16x8(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ float softplus(float z) {
  const float bz = BETA * z;
  return bz > 20.0f ? z : __logf(1.0f + __expf(bz)) * (1.0f / BETA);
}

// Eddy viscosity of the calling lane's cell.  Must be called by all 64 lanes
// of the wave (inactive cells pass any finite features and ignore the result).
// P points to the packed blob (LDS or global), x to 7 raw features.
__device__ __forceinline__ float nu_t(const Packed* P, const float (&x)[7]) {
  const int lane = __lane_id();
  const int h = lane >> 5;
  u16x8 xb;
#pragma unroll
  for (int e = 0; e < 7; ++e) xb[e] = bf16_bits((x[e] - P->mean[e]) * P->istd[e]);
  xb[7] = 0;
  u16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
  const u16x8 b0 = h == 0 ? xb : zero;
  const u16x8 b1 = h == 0 ? zero : xb;
  const u16x8 a1lo = *reinterpret_cast<const u16x8*>(P->a1lo[lane]);
  const u16x8 a1hi = *reinterpret_cast<const u16x8*>(P->a1hi[lane]);
  f32x16 c0, c1;
#pragma unroll
  for (int r = 0; r < 16; ++r) c0[r] = c1[r] = P->b1c[h][r];
  c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a1lo), as_bf16x8(b0), c0, 0, 0, 0);
  c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a1hi), as_bf16x8(b1), c1, 0, 0, 0);
  // ReLU, narrow to bf16: the layer-1 C tile becomes the layer-2 B operand
  u16x8 h00, h01, h10, h11;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    h00[e] = bf16_bits(fmaxf(c0[e], 0.0f));
    h01[e] = bf16_bits(fmaxf(c0[8 + e], 0.0f));
    h10[e] = bf16_bits(fmaxf(c1[e], 0.0f));
    h11[e] = bf16_bits(fmaxf(c1[8 + e], 0.0f));
  }
  const u16x8 a20 = *reinterpret_cast<const u16x8*>(P->a2[0][lane]);
  const u16x8 a21 = *reinterpret_cast<const u16x8*>(P->a2[1][lane]);
  f32x16 d0, d1;
#pragma unroll
  for (int r = 0; r < 16; ++r) d0[r] = d1[r] = P->b2c[h][r];
  d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a20), as_bf16x8(h00), d0, 0, 0, 0);
  d0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a21), as_bf16x8(h01), d0, 0, 0, 0);
  d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a20), as_bf16x8(h10), d1, 0, 0, 0);
  d1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a21), as_bf16x8(h11), d1, 0, 0, 0);
  float p0 = 0.0f, p1 = 0.0f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const float w = P->w3c[h][r];
    p0 += w * fmaxf(d0[r], 0.0f);
    p1 += w * fmaxf(d1[r], 0.0f);
  }
  p0 += __shfl_xor(p0, 32, 64);
  p1 += __shfl_xor(p1, 32, 64);
  return softplus((h == 0 ? p0 : p1) + P->b3);
}

}  // namespace mlp
}  // namespace tlbm
