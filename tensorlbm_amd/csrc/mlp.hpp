// Learned eddy-viscosity closure: 7 -> 32 -> 32 -> 1 MLP (ReLU hidden layers,
// Softplus(beta=10) output), the 3-D counterpart of the reference's
// EddyViscosityMLP (src/tensorlbm/ai/model.py:40-92).
//
// Inference inside the collision runs on the matrix cores: per wave (64 cells,
// one per lane) the two hidden layers are v_mfma_f32_32x32x16_f16 tiles (fp16
// operands, exact products, fp32 accumulation).  Layout (wave64, 32x32x16):
//   A[i][k] : lane i + 32*(k/8), element k%8
//   B[k][j] : lane j + 32*(k/8), element k%8
//   C[i][j] : lane j + 32*((i/4)%2), element (i%4) + 4*(i/8)
// Hidden units are rows (i) and cells are columns (j).  The normalised inputs
// enter as two-term fp16 splits v = hi + lo (hi = fp16(v), lo = fp16(v - hi),
// ~22 bits); layer 1 puts the hi parts at K 0..6 and the lo parts at K 8..14 of
// one tile per 32-cell block (one cross-half lane exchange supplies the partner
// lane's half), and the bias, split the same way, at K 7 and 15 against a
// constant 1, so the accumulator starts from zero.  The layer-1 C tile becomes the layer-2 B operand in registers:
// relu(c) rounded to fp16 (11 bits, packed convert + packed max, one
// instruction per value), W2's columns permuted at packing time to match the
// C layout.  The output layer is 16 local fp32 FMAs plus one cross-half
// shuffle.  Six matrix-core tiles per wave.  The operand blob (3.5 KB) is read
// straight from global memory, where it stays cache-resident across the launch.
//
// The bf16 form (nu_t_bf16, blob from tlbm_mlp_pack_bf16) runs the same tiles
// with v_mfma_f32_32x32x16_bf16: W1 and W2 are single bf16 values (8 bits);
// the normalised inputs and the layer-1 bias enter as two-term bf16 splits
// (~16 bits) in the same K slots; the layer-1 activations become two B
// operands, the bf16 value of relu(c) and the bf16 value of what it leaves out,
// taken against the same W2 operand (four tiles per block for layer 2).  The
// output layer is the same fp32 step.  Ten matrix-core tiles per wave.
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

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Matrix-core operand blob, built once per weight update by tlbm_mlp_pack.
struct Packed {
  u16x8 a1[64];     // layer-1 A (fp16 bits): W1 row (lane % 32), inputs 0..6 (both K halves),
                    // element 7: b1 hi (lanes < 32) / lo (lanes >= 32)
  u16x8 a2[2][64];  // layer-2 A (fp16 bits), K-steps 0 and 1, columns in C-layout order
  float b1c[2][16];       // layer-1 bias, C layout per lane half (reference copy; the tile takes it via K)
  float b2c[2][16];       // layer-2 bias
  float w3c[2][16];       // output weights
  float b3;
  float mean[7];
  float istd[7];
  float pad;
};
static_assert(sizeof(Packed) == 3520, "packed MLP blob layout");
// the blob as staged in a workgroup's LDS
typedef __attribute__((address_space(3))) const Packed* LdsPacked;

// hidden unit held by C-tile element r of a lane in half h (= lane / 32)
__host__ __device__ constexpr int c_row(int h, int r) { return 4 * h + (r % 4) + 8 * (r / 4); }
// W2 column used by layer-2 K-step s, lane half h, element e
__host__ __device__ constexpr int w2_col(int s, int h, int e) { return 4 * h + (e % 4) + 16 * s + 8 * (e / 4); }


__device__ __forceinline__ f16x8 as_f16x8(u16x8 v) { return __builtin_bit_cast(f16x8, v); }

// log(1 + e^(beta z)) / beta on the hardware base-2 exponential and logarithm
// (v_exp_f32 / v_log_f32, about 1 ulp each): 1 + e^(beta z) >= 1 is never
// denormal, so neither needs the range scaling of the library functions
__device__ __forceinline__ float softplus(float z) {
  const float bz = BETA * z;
  const float e = __builtin_amdgcn_exp2f(bz * 1.4426950408889634f);
  return bz > 20.0f ? z : __builtin_amdgcn_logf(1.0f + e) * (0.6931471805599453f / BETA);
}

// max(v, 0) as one integer max on the bit pattern (negative floats are negative
// integers); exact for every non-NaN v
__device__ __forceinline__ float relu(float v) {
  return __builtin_bit_cast(float, __builtin_elementwise_max(__builtin_bit_cast(int, v), 0));
}

// the partner lane's (lane ^ 32) operand, for either 16-bit operand type
template <typename V>
__device__ __forceinline__ V xor32(V v) {
  u32x4 u = __builtin_bit_cast(u32x4, v);
#pragma unroll
  for (int i = 0; i < 4; ++i) u[i] = uint32_t(__shfl_xor(int(u[i]), 32, 64));
  return __builtin_bit_cast(V, u);
}

// relu(c[8s .. 8s+7]) rounded to fp16 (round to nearest even), as a layer-2 B operand
__device__ __forceinline__ f16x8 act16(const f32x16& c, int s) {
  f16x8 b;
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    const f32x2 p = {c[8 * s + e], c[8 * s + e + 1]};
    f16x2 hh = __builtin_convertvector(p, f16x2);
    hh = __builtin_elementwise_max(hh, (f16x2){_Float16(0), _Float16(0)});
    b[e] = hh[0];
    b[e + 1] = hh[1];
  }
  return b;
}

// layer 2 + output for one 32-cell block whose layer-1 tile is c
template <class PP>
__device__ __forceinline__ float block_out(PP P, int h, const f32x16& c, f16x8 a20, f16x8 a21) {
  f32x16 d;
#pragma unroll
  for (int r = 0; r < 16; ++r) d[r] = P->b2c[h][r];
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a20, act16(c, 0), d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a21, act16(c, 1), d, 0, 0, 0);
  float p = -0.0f;  // -0: the first fma folds into a multiply
#pragma unroll
  for (int r = 0; r < 16; ++r) p += P->w3c[h][r] * relu(d[r]);
  return p + __shfl_xor(p, 32, 64);
}

// Eddy viscosity of the calling lane's cell.  Must be called by all 64 lanes
// of the wave (inactive cells pass any finite features and ignore the result).
// P points to the packed blob (a global or an LDS pointer), x to 7 raw features.
template <class PP>
__device__ __forceinline__ float nu_t(PP P, const float (&x)[7]) {
  const int lane = __lane_id();
  const int h = lane >> 5;
  f16x8 hi, lo;
#pragma unroll
  for (int e = 0; e < 7; ++e) {
    const float v = (x[e] - P->mean[e]) * P->istd[e];
    hi[e] = static_cast<_Float16>(v);
    lo[e] = static_cast<_Float16>(v - static_cast<float>(hi[e]));
  }
  // K 7 and 15 carry the layer-1 bias (hi and lo parts in A) against a constant 1
  hi[7] = lo[7] = _Float16(1);
  // block 0 (cells of lanes 0..31): K 0..7 = own hi on lanes < 32, K 8..15 = the
  // partner's lo on lanes >= 32; block 1 (cells of lanes 32..63) the other way round
  const f16x8 got = xor32(h == 0 ? lo : hi);
  const f16x8 b0 = h == 0 ? hi : got;
  const f16x8 b1 = h == 0 ? got : lo;
  const f16x8 a1 = as_f16x8(P->a1[lane]);
  const f16x8 a20 = as_f16x8(P->a2[0][lane]);
  const f16x8 a21 = as_f16x8(P->a2[1][lane]);
  // one 32-cell block at a time keeps a single accumulator tile pair live
  const f32x16 zero = {};
  f32x16 c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, b0, zero, 0, 0, 0);
  const float p0 = block_out(P, h, c, a20, a21);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, b1, zero, 0, 0, 0);
  const float p1 = block_out(P, h, c, a20, a21);
  return softplus((h == 0 ? p0 : p1) + P->b3);
}

// ---- bf16 operands ----------------------------------------------------------

__device__ __forceinline__ bf16x8 as_bf16x8(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

// relu(c[8s .. 8s+7]) as a two-term bf16 split: hi = bf16(v), lo = bf16(v - hi), both
// rounded to nearest even; hi + lo carries ~16 bits of v
__device__ __forceinline__ void act_bf16(const f32x16& c, int s, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float v = relu(c[8 * s + e]);
    const __bf16 h = static_cast<__bf16>(v);
    hi[e] = h;
    lo[e] = static_cast<__bf16>(v - static_cast<float>(h));
  }
}

// layer 2 (hi and lo activations against the same W2 tile) + output for one 32-cell block
template <class PP>
__device__ __forceinline__ float block_out_bf16(PP P, int h, const f32x16& c, bf16x8 a20, bf16x8 a21) {
  f32x16 d;
#pragma unroll
  for (int r = 0; r < 16; ++r) d[r] = P->b2c[h][r];
  bf16x8 bh, bl;
  act_bf16(c, 0, bh, bl);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a20, bh, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a20, bl, d, 0, 0, 0);
  act_bf16(c, 1, bh, bl);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a21, bh, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a21, bl, d, 0, 0, 0);
  float p = -0.0f;  // -0: the first fma folds into a multiply
#pragma unroll
  for (int r = 0; r < 16; ++r) p += P->w3c[h][r] * relu(d[r]);
  return p + __shfl_xor(p, 32, 64);
}

// nu_t of the bf16 blob (tlbm_mlp_pack_bf16); same calling rules as nu_t
template <class PP>
__device__ __forceinline__ float nu_t_bf16(PP P, const float (&x)[7]) {
  const int lane = __lane_id();
  const int h = lane >> 5;
  bf16x8 hi, lo;
#pragma unroll
  for (int e = 0; e < 7; ++e) {
    const float v = (x[e] - P->mean[e]) * P->istd[e];
    hi[e] = static_cast<__bf16>(v);
    lo[e] = static_cast<__bf16>(v - static_cast<float>(hi[e]));
  }
  hi[7] = lo[7] = static_cast<__bf16>(1.0f);
  const bf16x8 got = xor32(h == 0 ? lo : hi);
  const bf16x8 b0 = h == 0 ? hi : got;
  const bf16x8 b1 = h == 0 ? got : lo;
  const bf16x8 a1 = as_bf16x8(P->a1[lane]);
  const bf16x8 a20 = as_bf16x8(P->a2[0][lane]);
  const bf16x8 a21 = as_bf16x8(P->a2[1][lane]);
  const f32x16 zero = {};
  f32x16 c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, zero, 0, 0, 0);
  const float p0 = block_out_bf16(P, h, c, a20, a21);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, zero, 0, 0, 0);
  const float p1 = block_out_bf16(P, h, c, a20, a21);
  return softplus((h == 0 ? p0 : p1) + P->b3);
}

}  // namespace mlp
}  // namespace tlbm
