// Learned eddy-viscosity closure: packing, training and data kernels.
//
// Training is deterministic: tlbm_mlp_grad writes one partial gradient vector
// per 64-sample workgroup, tlbm_mlp_adam sums the partials in a fixed order
// and applies Adam (torch.optim.Adam update rule, as in the reference's
// train_eddy_viscosity_model, src/tensorlbm/ai/train.py:382-435).
//
// Data: the DNS populations are block-averaged by an integer factor
// (the 3-D form of ai/pipeline.py:93-107 _coarsen_mean) into rho, u_i and
// <u_i u_j>; the samples are the resolved strain (periodic central
// differences, ai/dataset.py:41-72 in 3-D) and the sub-grid eddy viscosity
// nu = max(0, -tau_ij S_ij / (2 S_kl S_kl)), tau_ij = <u_i u_j> - <u_i><u_j>.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>

#include "../../include/tlbm_hip.h"
#include "collide.hpp"
#include "mlp.hpp"
#include "store.hpp"

namespace tlbm {

using namespace mlp;

constexpr int GRAD_BS = 64;  // samples per workgroup in the gradient kernel

// ---------------------------------------------------------------------------
// packing: fp32 parameters -> matrix-core operand blob
// ---------------------------------------------------------------------------

// fp16 bits of v, round to nearest even
__device__ __forceinline__ uint16_t f16_bits(float v) { return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v)); }

// bf16 bits of v, round to nearest even
__device__ __forceinline__ uint16_t bf16_bits(float v) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v)); }

// operand bits of v in the blob's type, and v rounded to that type
template <bool BF16>
__device__ __forceinline__ uint16_t op_bits(float v) {
  return BF16 ? bf16_bits(v) : f16_bits(v);
}
template <bool BF16>
__device__ __forceinline__ float op_round(float v) {
  return BF16 ? static_cast<float>(static_cast<__bf16>(v)) : static_cast<float>(static_cast<_Float16>(v));
}

template <bool BF16>
__global__ __launch_bounds__(64) void k_mlp_pack(const float* __restrict__ p, Packed* __restrict__ out) {
  const int l = threadIdx.x;
  const int h = l >> 5, i = l & 31;
  for (int e = 0; e < 8; ++e) {
    if (e < NIN) {
      out->a1[l][e] = op_bits<BF16>(p[OW1 + i * NIN + e]);
    } else {  // bias at K 7 (hi, lanes < 32) and K 15 (lo, lanes >= 32)
      const float b = p[OB1 + i];
      out->a1[l][e] = op_bits<BF16>(h == 0 ? b : b - op_round<BF16>(b));
    }
    for (int s = 0; s < 2; ++s) out->a2[s][l][e] = op_bits<BF16>(p[OW2 + i * NH + w2_col(s, h, e)]);
  }
  if (l < 32) {
    const int hh = l >> 4, r = l & 15;
    out->b1c[hh][r] = p[OB1 + c_row(hh, r)];
    out->b2c[hh][r] = p[OB2 + c_row(hh, r)];
    out->w3c[hh][r] = p[OW3 + c_row(hh, r)];
  }
  if (l == 0) {
    out->b3 = p[OB3];
    for (int e = 0; e < NIN; ++e) {
      out->mean[e] = p[OMEAN + e];
      out->istd[e] = p[OISTD + e];
    }
    out->pad = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// fp32 forward (prediction) and gradient
// ---------------------------------------------------------------------------

struct Fwd {
  float xh[NIN], z1[NH], z2[NH], z3, y;
};

__device__ __forceinline__ void forward32(const float* __restrict__ p, const float* x, Fwd& r) {
#pragma unroll
  for (int e = 0; e < NIN; ++e) r.xh[e] = (x[e] - p[OMEAN + e]) * p[OISTD + e];
  for (int i = 0; i < NH; ++i) {
    float a = p[OB1 + i];
#pragma unroll
    for (int e = 0; e < NIN; ++e) a += p[OW1 + i * NIN + e] * r.xh[e];
    r.z1[i] = a;
  }
  for (int i = 0; i < NH; ++i) {
    float a = p[OB2 + i];
    for (int j = 0; j < NH; ++j) a += p[OW2 + i * NH + j] * fmaxf(r.z1[j], 0.0f);
    r.z2[i] = a;
  }
  float a = p[OB3];
  for (int i = 0; i < NH; ++i) a += p[OW3 + i] * fmaxf(r.z2[i], 0.0f);
  r.z3 = a;
  r.y = softplus(a);
}

__global__ __launch_bounds__(256) void k_mlp_predict(const float* __restrict__ p, const float* __restrict__ x,
                                                     uint64_t n, float* __restrict__ y) {
  const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float xs[NIN];
#pragma unroll
  for (int e = 0; e < NIN; ++e) xs[e] = x[s * NIN + e];
  Fwd r;
  forward32(p, xs, r);
  y[s] = r.y;
}

// Each workgroup: 64 samples (one per lane of wave 0) run forward + backward,
// park activations and deltas in LDS, then 256 lanes reduce the 1345 parameter
// gradients (and the loss) over the 64 samples into partials[block].
__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ p, const float* __restrict__ x,
                                                  const float* __restrict__ t, uint64_t n, float inv_n,
                                                  float* __restrict__ partials) {
  __shared__ float s_xh[GRAD_BS][8];
  __shared__ float s_h1[GRAD_BS][NH];
  __shared__ float s_d1[GRAD_BS][NH];
  __shared__ float s_h2[GRAD_BS][NH];
  __shared__ float s_d2[GRAD_BS][NH];
  __shared__ float s_d3[GRAD_BS];
  __shared__ float s_loss[GRAD_BS];
  const int tid = threadIdx.x;
  if (tid < GRAD_BS) {
    const uint64_t s = uint64_t(blockIdx.x) * GRAD_BS + tid;
    const bool ok = s < n;
    float xs[NIN];
#pragma unroll
    for (int e = 0; e < NIN; ++e) xs[e] = ok ? x[s * NIN + e] : 0.0f;
    Fwd r;
    forward32(p, xs, r);
    const float res = ok ? r.y - t[s] : 0.0f;
    s_loss[tid] = res * res * inv_n;
    // dL/dz3 = 2 (y - t)/N * softplus'(z3): sigmoid(beta z3), 1 where softplus returns z3 (beta z3 > 20)
    const float bz = fmaxf(BETA * r.z3, -60.0f);
    const float d3 = 2.0f * res * inv_n * (bz > 20.0f ? 1.0f : 1.0f / (1.0f + __expf(-bz)));
    s_d3[tid] = d3;
    float d2[NH];
    for (int i = 0; i < NH; ++i) {
      d2[i] = r.z2[i] > 0.0f ? d3 * p[OW3 + i] : 0.0f;
      s_d2[tid][i] = d2[i];
      s_h2[tid][i] = fmaxf(r.z2[i], 0.0f);
    }
    for (int j = 0; j < NH; ++j) {
      float a = 0.0f;
      for (int i = 0; i < NH; ++i) a += d2[i] * p[OW2 + i * NH + j];
      s_d1[tid][j] = r.z1[j] > 0.0f ? a : 0.0f;
      s_h1[tid][j] = fmaxf(r.z1[j], 0.0f);
    }
#pragma unroll
    for (int e = 0; e < NIN; ++e) s_xh[tid][e] = r.xh[e];
    s_xh[tid][7] = 0.0f;
  }
  __syncthreads();
  float* out = partials + uint64_t(blockIdx.x) * (NPARAM + 1);
  for (int k = tid; k <= NPARAM; k += blockDim.x) {
    float g = 0.0f;
    if (k < OB1) {
      const int i = k / NIN, e = k % NIN;
      for (int s = 0; s < GRAD_BS; ++s) g += s_d1[s][i] * s_xh[s][e];
    } else if (k < OW2) {
      const int i = k - OB1;
      for (int s = 0; s < GRAD_BS; ++s) g += s_d1[s][i];
    } else if (k < OB2) {
      const int i = (k - OW2) / NH, j = (k - OW2) % NH;
      for (int s = 0; s < GRAD_BS; ++s) g += s_d2[s][i] * s_h1[s][j];
    } else if (k < OW3) {
      const int i = k - OB2;
      for (int s = 0; s < GRAD_BS; ++s) g += s_d2[s][i];
    } else if (k < OB3) {
      const int i = k - OW3;
      for (int s = 0; s < GRAD_BS; ++s) g += s_d3[s] * s_h2[s][i];
    } else if (k == OB3) {
      for (int s = 0; s < GRAD_BS; ++s) g += s_d3[s];
    } else {
      for (int s = 0; s < GRAD_BS; ++s) g += s_loss[s];
    }
    out[k] = g;
  }
}

// One workgroup: sum the per-workgroup partials in block order, clip the
// global gradient norm (torch.nn.utils.clip_grad_norm_ rule: scale by
// min(1, clip / (|g| + 1e-6)), ai/train.py:401-405), then one Adam step.
constexpr int ADAM_THREADS = 1024;
static_assert(2 * ADAM_THREADS > NPARAM, "two entries per thread cover the gradient and the loss");

__global__ __launch_bounds__(ADAM_THREADS) void k_mlp_adam(float* __restrict__ p, const float* __restrict__ partials,
                                                           int nblk, float* __restrict__ m, float* __restrict__ v,
                                                           float lr, float b1, float b2, float eps, float bc1,
                                                           float bc2, float clip, float* __restrict__ loss_out,
                                                           float* __restrict__ grad_out) {
  __shared__ float s_red[ADAM_THREADS];
  const int tid = threadIdx.x;
  float g[2];
  float sq = 0.0f;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int k = tid + r * ADAM_THREADS;
    float a = 0.0f;
    if (k <= NPARAM)
      for (int b = 0; b < nblk; ++b) a += partials[uint64_t(b) * (NPARAM + 1) + k];
    g[r] = a;
    if (k < NPARAM) sq += a * a;
  }
  s_red[tid] = sq;
  __syncthreads();
  for (int off = ADAM_THREADS / 2; off > 0; off >>= 1) {
    if (tid < off) s_red[tid] += s_red[tid + off];
    __syncthreads();
  }
  const float norm = sqrtf(s_red[0]);
  const float scale = clip > 0.0f ? fminf(1.0f, clip / (norm + 1e-6f)) : 1.0f;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int k = tid + r * ADAM_THREADS;
    if (k == NPARAM && loss_out) *loss_out = g[r];
    if (k >= NPARAM) continue;
    if (grad_out) grad_out[k] = g[r];
    if (lr == 0.0f) continue;
    const float gk = g[r] * scale;
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
  }
}

// Gradient partials of tlbm_mlp_train_step_noalloc: up to NOALLOC_MAX_BLOCKS workgroups of
// GRAD_BS samples, (NPARAM + 1) floats each (2.76 MB), one copy per GPU for the life of the
// process.  k_mlp_grad writes every entry of the blocks a call launches before k_mlp_adam
// reads them, so nothing is cleared between calls.
constexpr int NOALLOC_MAX_BLOCKS = TLBM_MLP_NOALLOC_MAX_SAMPLES / GRAD_BS;
static_assert(NOALLOC_MAX_BLOCKS * GRAD_BS == TLBM_MLP_NOALLOC_MAX_SAMPLES, "whole blocks");
static __device__ __align__(256) float scratch[NOALLOC_MAX_BLOCKS * (NPARAM + 1)];

// ---------------------------------------------------------------------------
// DNS -> LES data
// ---------------------------------------------------------------------------

struct Lay {
  int64_t sq, sz, sy;
  int nx, ny, nz;
};

// out: 10 fields of the coarse grid: rho, u_x, u_y, u_z, <uxux> <uyuy> <uzuz> <uxuy> <uxuz> <uyuz>
template <class L, int S>
__global__ __launch_bounds__(256) void k_coarsen(const void* __restrict__ fv, Lay lay, int r, float* __restrict__ out) {
  using ST = Store<S>;
  const typename ST::T* f = static_cast<const typename ST::T*>(fv);
  const int cx = lay.nx / r, cy = lay.ny / r, cz = lay.nz / r;
  const uint64_t nc = uint64_t(cx) * cy * cz;
  const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int X = int(c % cx), Y = int((c / cx) % cy), Z = int(c / (uint64_t(cx) * cy));
  double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int dz = 0; dz < r; ++dz)
    for (int dy = 0; dy < r; ++dy)
      for (int dx = 0; dx < r; ++dx) {
        const int64_t base = int64_t(Z * r + dz) * lay.sz + int64_t(Y * r + dy) * lay.sy + int64_t(X * r + dx);
        typename ST::C v[L::Q];
#pragma unroll
        for (int q = 0; q < L::Q; ++q) v[q] = ST::template get<L>(f[base + int64_t(q) * lay.sq], q);
        const Macro<typename ST::C> mc = macro<L>(v);
        const double ux = mc.ux, uy = mc.uy, uz = mc.uz;
        acc[0] += mc.rho;
        acc[1] += ux;
        acc[2] += uy;
        acc[3] += uz;
        acc[4] += ux * ux;
        acc[5] += uy * uy;
        acc[6] += uz * uz;
        acc[7] += ux * uy;
        acc[8] += ux * uz;
        acc[9] += uy * uz;
      }
  const double inv = 1.0 / double(r * r * r);
  for (int k = 0; k < 10; ++k) out[uint64_t(k) * nc + c] = float(acc[k] * inv);
}

__global__ __launch_bounds__(256) void k_sgs_samples(const float* __restrict__ fld, int nx, int ny, int nz,
                                                     float* __restrict__ feats, float* __restrict__ labels) {
  const uint64_t nc = uint64_t(nx) * ny * nz;
  const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int x = int(c % nx), y = int((c / nx) % ny), z = int(c / (uint64_t(nx) * ny));
  const int xp = x == nx - 1 ? 0 : x + 1, xm = x == 0 ? nx - 1 : x - 1;
  const int yp = y == ny - 1 ? 0 : y + 1, ym = y == 0 ? ny - 1 : y - 1;
  const int zp = z == nz - 1 ? 0 : z + 1, zm = z == 0 ? nz - 1 : z - 1;
  auto at = [&](int k, int zz, int yy, int xx) {
    return double(fld[uint64_t(k) * nc + (uint64_t(zz) * ny + yy) * nx + xx]);
  };
  double g[3][3];  // g[i][j] = d u_i / d x_j
  for (int i = 0; i < 3; ++i) {
    g[i][0] = 0.5 * (at(1 + i, z, y, xp) - at(1 + i, z, y, xm));
    g[i][1] = 0.5 * (at(1 + i, z, yp, x) - at(1 + i, z, ym, x));
    g[i][2] = 0.5 * (at(1 + i, zp, y, x) - at(1 + i, zm, y, x));
  }
  const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
  const double ux = at(1, z, y, x), uy = at(2, z, y, x), uz = at(3, z, y, x);
  const double um[3] = {ux, uy, uz};
  double ss = 0.0, ts = 0.0;
  for (int k = 0; k < 6; ++k) {
    const int i = ci[k], j = cj[k];
    const double s = 0.5 * (g[i][j] + g[j][i]);
    const double tau = at(4 + k, z, y, x) - um[i] * um[j];
    const double w = i == j ? 1.0 : 2.0;
    ss += w * s * s;
    ts += w * tau * s;
    feats[c * NIN + k] = float(s);
  }
  feats[c * NIN + 6] = float(sqrt(2.0 * ss));
  const double nu = -ts / (2.0 * (ss > 1e-30 ? ss : 1e-30));
  labels[c] = float(nu > 0.0 ? nu : 0.0);
}

}  // namespace tlbm

// ===========================================================================
// C ABI
// ===========================================================================

using namespace tlbm;

namespace tlbm {
int set_error(int code, const char* msg);  // tlbm.hip: shared tlbm_last_error() channel
}

namespace {

int cfail(int code, const char* msg) { return tlbm::set_error(code, msg); }

int claunch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    return tlbm::set_error(TLBM_ERR_HIP, buf);
  }
  return TLBM_OK;
}

// Device address of `scratch` on the current device.  Host code must never pass the array's
// name to a kernel: on the host it is a stand-in, not the device copy's address.  Looked up once
// per device (each GPU holds its own copy at its own address) and kept.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float* g_scratch[MAX_DEVICES] = {};

int scratch_ptr(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return cfail(TLBM_ERR_HIP, "hipGetDevice failed");
  if (dev < 0 || dev >= MAX_DEVICES) return cfail(TLBM_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(tlbm::scratch));
    if (e != hipSuccess || p == nullptr) {
      char buf[256];
      snprintf(buf, sizeof(buf), "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return tlbm::set_error(TLBM_ERR_HIP, buf);
    }
    g_scratch[dev] = static_cast<float*>(p);
  }
  *out = g_scratch[dev];
  return TLBM_OK;
}

}  // namespace

extern "C" {

int tlbm_mlp_sizes(int64_t* nparam, int64_t* ntotal, int64_t* packed_bytes, int64_t* batch_block) {
  if (nparam) *nparam = mlp::NPARAM;
  if (ntotal) *ntotal = mlp::NTOTAL;
  if (packed_bytes) *packed_bytes = int64_t(sizeof(mlp::Packed));
  if (batch_block) *batch_block = GRAD_BS;
  return TLBM_OK;
}

int tlbm_mlp_pack(const float* params, void* packed, tlbm_stream_t stream) {
  if (!params || !packed) return cfail(TLBM_ERR_INVALID, "null pointer");
  hipLaunchKernelGGL(k_mlp_pack<false>, dim3(1), dim3(64), 0, (hipStream_t)stream, params, (mlp::Packed*)packed);
  return claunch("k_mlp_pack");
}

int tlbm_mlp_pack_bf16(const float* params, void* packed, tlbm_stream_t stream) {
  if (!params || !packed) return cfail(TLBM_ERR_INVALID, "null pointer");
  hipLaunchKernelGGL(k_mlp_pack<true>, dim3(1), dim3(64), 0, (hipStream_t)stream, params, (mlp::Packed*)packed);
  return claunch("k_mlp_pack");
}

int tlbm_mlp_predict(const float* params, const float* x, int64_t n, float* y, tlbm_stream_t stream) {
  if (n < 0) return cfail(TLBM_ERR_SHAPE, "negative sample count");
  if (n == 0) return TLBM_OK;
  if (!params || !x || !y) return cfail(TLBM_ERR_INVALID, "null pointer");
  const uint32_t nb = uint32_t((uint64_t(n) + 255) / 256);
  hipLaunchKernelGGL(k_mlp_predict, dim3(nb), dim3(256), 0, (hipStream_t)stream, params, x, uint64_t(n), y);
  return claunch("k_mlp_predict");
}

int tlbm_mlp_grad(const float* params, const float* x, const float* t, int64_t n, float* partials,
                  tlbm_stream_t stream) {
  if (n <= 0) return cfail(TLBM_ERR_SHAPE, "need at least one sample");
  if (!params || !x || !t || !partials) return cfail(TLBM_ERR_INVALID, "null pointer");
  const uint64_t nb = (uint64_t(n) + GRAD_BS - 1) / GRAD_BS;
  if (nb >= (uint64_t(1) << 31)) return cfail(TLBM_ERR_SHAPE, "batch too large");
  hipLaunchKernelGGL(k_mlp_grad, dim3(uint32_t(nb)), dim3(256), 0, (hipStream_t)stream, params, x, t, uint64_t(n),
                     1.0f / float(n), partials);
  return claunch("k_mlp_grad");
}

int tlbm_mlp_adam(float* params, const float* partials, int nblocks, float* m, float* v, int step, float lr,
                  float beta1, float beta2, float eps, float clip_norm, float* loss_out, float* grad_out,
                  tlbm_stream_t stream) {
  if (nblocks <= 0 || step < 1) return cfail(TLBM_ERR_INVALID, "need nblocks > 0 and step >= 1");
  if (!params || !partials || (lr != 0.0f && (!m || !v))) return cfail(TLBM_ERR_INVALID, "null pointer");
  const float bc1 = 1.0f - powf(beta1, float(step));
  const float bc2 = 1.0f - powf(beta2, float(step));
  hipLaunchKernelGGL(k_mlp_adam, dim3(1), dim3(ADAM_THREADS), 0, (hipStream_t)stream, params, partials, nblocks, m,
                     v, lr, beta1, beta2, eps, bc1, bc2, clip_norm, loss_out, grad_out);
  return claunch("k_mlp_adam");
}

int tlbm_mlp_train_step_noalloc(float* params, const float* x, const float* t, int64_t n, float* m, float* v,
                                int step, float lr, float beta1, float beta2, float eps, float clip_norm,
                                float* loss_out, float* grad_out, tlbm_stream_t stream) {
  if (n <= 0) return cfail(TLBM_ERR_SHAPE, "need at least one sample");
  if (n > TLBM_MLP_NOALLOC_MAX_SAMPLES) {
    char buf[160];
    snprintf(buf, sizeof(buf), "%lld samples exceed the no-allocation limit of %d per call", (long long)n,
             TLBM_MLP_NOALLOC_MAX_SAMPLES);
    return cfail(TLBM_ERR_SHAPE, buf);
  }
  if (step < 1) return cfail(TLBM_ERR_INVALID, "need step >= 1");
  if (!params || !x || !t || (lr != 0.0f && (!m || !v))) return cfail(TLBM_ERR_INVALID, "null pointer");
  float* partials = nullptr;
  const int rc = scratch_ptr(&partials);
  if (rc != TLBM_OK) return rc;
  const int nb = int((n + GRAD_BS - 1) / GRAD_BS);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_grad, dim3(uint32_t(nb)), dim3(256), 0, s, params, x, t, uint64_t(n), 1.0f / float(n),
                     partials);
  const int rg = claunch("k_mlp_grad");
  if (rg != TLBM_OK) return rg;
  const float bc1 = 1.0f - powf(beta1, float(step));
  const float bc2 = 1.0f - powf(beta2, float(step));
  hipLaunchKernelGGL(k_mlp_adam, dim3(1), dim3(ADAM_THREADS), 0, s, params, partials, nb, m, v, lr, beta1, beta2,
                     eps, bc1, bc2, clip_norm, loss_out, grad_out);
  return claunch("k_mlp_adam");
}

int tlbm_coarsen_moments(int lattice, int store, const void* f, const int64_t* strides, int nz, int ny, int nx,
                         int factor, float* out, tlbm_stream_t stream) {
  if (!f || !out) return cfail(TLBM_ERR_INVALID, "null pointer");
  if (factor < 1 || nx % factor || ny % factor || nz % factor)
    return cfail(TLBM_ERR_SHAPE, "grid dims must be multiples of the coarsening factor");
  if (lattice != TLBM_D3Q19 && lattice != TLBM_D3Q27) return cfail(TLBM_ERR_UNSUPPORTED, "3-D lattices only");
  const int64_t N = int64_t(nz) * ny * nx;
  Lay lay{strides ? strides[0] : N, strides ? strides[1] : int64_t(ny) * nx, strides ? strides[2] : nx, nx, ny, nz};
  const uint64_t nc = uint64_t(nx / factor) * (ny / factor) * (nz / factor);
  const uint32_t nb = uint32_t((nc + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
#define TLBM_COARSEN(LT, SK) \
  hipLaunchKernelGGL((k_coarsen<LT, SK>), dim3(nb), dim3(256), 0, s, f, lay, factor, out)
  if (lattice == TLBM_D3Q19) {
    switch (store) {
      case TLBM_STORE_F32: TLBM_COARSEN(D3Q19, ST_F32); break;
      case TLBM_STORE_F64: TLBM_COARSEN(D3Q19, ST_F64); break;
      case TLBM_STORE_F16_DEV: TLBM_COARSEN(D3Q19, ST_F16_DEV); break;
      default: return cfail(TLBM_ERR_DTYPE, "coarsening takes fp32, fp64 or fp16-deviation populations");
    }
  } else {
    switch (store) {
      case TLBM_STORE_F32: TLBM_COARSEN(D3Q27, ST_F32); break;
      case TLBM_STORE_F64: TLBM_COARSEN(D3Q27, ST_F64); break;
      default: return cfail(TLBM_ERR_DTYPE, "coarsening takes fp32 or fp64 D3Q27 populations");
    }
  }
#undef TLBM_COARSEN
  return claunch("k_coarsen");
}

int tlbm_sgs_samples(const float* fields, int nz, int ny, int nx, float* feats, float* labels, tlbm_stream_t stream) {
  if (!fields || !feats || !labels) return cfail(TLBM_ERR_INVALID, "null pointer");
  if (nx < 3 || ny < 3 || nz < 3) return cfail(TLBM_ERR_SHAPE, "coarse grid needs at least 3 cells per axis");
  const uint64_t nc = uint64_t(nx) * ny * nz;
  hipLaunchKernelGGL(k_sgs_samples, dim3(uint32_t((nc + 255) / 256)), dim3(256), 0, (hipStream_t)stream, fields, nx,
                     ny, nz, feats, labels);
  return claunch("k_sgs_samples");
}

}  // extern "C"
