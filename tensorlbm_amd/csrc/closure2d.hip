// The reference's 2-D learned closure on the GPU: a small MLP (in -> H -> ... -> H -> 1, tanh,
// relu or gelu hidden activations, Softplus(beta = 10) output; ai/model.py:19-92) mapping the
// strain rate of a D2Q9 velocity field (S_xx, S_yy, S_xy; ai/dataset.py:41-72) to an eddy
// viscosity, its training (MSE, clipped gradient norm, Adam; ai/train.py:310-464) and the LES
// collision that uses it (collide_ai_les_bgk, ai/inference.py:53-92).
//
// Parameters are one fp32 vector in nn.Sequential order: for each hidden layer W [H][in] and
// b [H], then the output w [H] and b [1], then the input normalisation mean [in] and std [in]
// (the network sees (x - mean) / std).  Hidden widths 8, 16, 32 and 64 are built (template H);
// the depth (1..4 hidden layers), the input count (1..8) and the activation are run-time.
//
// Training is deterministic: a workgroup runs 64 samples forward and backward (one per lane
// of wave 0), parks activations and deltas in LDS and reduces every parameter's gradient over
// its samples in a fixed order into one partial vector; the Adam kernel sums the partials in
// block order.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <type_traits>

#include "../../include/tlbm_hip.h"
#include "lattice.hpp"

namespace tlbm {

int set_error(int code, const char* msg);  // tlbm.hip: shared tlbm_last_error() channel

namespace mlp2d {

constexpr int MAX_IN = 8;
constexpr int MAX_LAYERS = 4;
constexpr int BS = 64;  // samples per gradient workgroup
constexpr float BETA = 10.0f;

struct Arch {
  int in, hidden, layers, act;  // act: TLBM_ACT_*
};

__host__ __device__ inline int64_t nparam(const Arch& a) {
  int64_t n = int64_t(a.hidden) * (a.in + 1);
  n += int64_t(a.layers - 1) * a.hidden * (a.hidden + 1);
  return n + a.hidden + 1;
}

__device__ __forceinline__ float act(int kind, float z) {
  if (kind == TLBM_ACT_TANH) return tanhf(z);
  if (kind == TLBM_ACT_RELU) return z > 0.0f ? z : 0.0f;
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));  // exact (erf) GELU
}

// d act / dz at pre-activation z
__device__ __forceinline__ float dact(int kind, float z) {
  if (kind == TLBM_ACT_TANH) {
    const float t = tanhf(z);
    return 1.0f - t * t;
  }
  if (kind == TLBM_ACT_RELU) return z > 0.0f ? 1.0f : 0.0f;
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  return cdf + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Softplus(beta) with torch's threshold: z where beta z > 20
__device__ __forceinline__ float softplus(float z) {
  const float bz = BETA * z;
  return bz > 20.0f ? z : log1pf(expf(bz)) / BETA;
}

// forward pass of one sample; z[l][j] pre-activations are written to zs (stride H per layer)
// when it is non-null; returns the output pre-activation
template <int H>
__device__ float forward(const float* __restrict__ p, const Arch& a, const float* xin, float* zs) {
  float h[H], z[H];
  const float* mean = p + nparam(a);
  const float* sd = mean + a.in;
  float x[MAX_IN];
#pragma unroll
  for (int e = 0; e < MAX_IN; ++e) x[e] = e < a.in ? (xin[e] - mean[e]) / sd[e] : 0.0f;
  const float* w = p;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float s = w[H * a.in + j];
    for (int e = 0; e < a.in; ++e) s = fmaf(w[j * a.in + e], x[e], s);
    z[j] = s;
  }
  w += H * (a.in + 1);
  for (int l = 0;; ++l) {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      if (zs) zs[l * H + j] = z[j];
      h[j] = act(a.act, z[j]);
    }
    if (l + 1 == a.layers) break;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      float s = w[H * H + j];
#pragma unroll
      for (int i = 0; i < H; ++i) s = fmaf(w[j * H + i], h[i], s);
      z[j] = s;
    }
    w += H * (H + 1);
  }
  float s = w[H];
#pragma unroll
  for (int i = 0; i < H; ++i) s = fmaf(w[i], h[i], s);
  return s;
}

template <int H>
__global__ __launch_bounds__(256) void k_forward(const float* __restrict__ p, Arch a, const float* __restrict__ x,
                                                 uint64_t n, float* __restrict__ y) {
  const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  y[s] = softplus(forward<H>(p, a, x + s * a.in, nullptr));
}

// periodic central-difference strain of u at (y, x) (ai/dataset.py:41-72)
__device__ __forceinline__ void strain_at(const float* ux, const float* uy, int ny, int nx, int y, int x,
                                          float (&s)[3]) {
  const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
  const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
  const float dudx = 0.5f * (ux[y * nx + xp] - ux[y * nx + xm]);
  const float dudy = 0.5f * (ux[yp * nx + x] - ux[ym * nx + x]);
  const float dvdx = 0.5f * (uy[y * nx + xp] - uy[y * nx + xm]);
  const float dvdy = 0.5f * (uy[yp * nx + x] - uy[ym * nx + x]);
  s[0] = dudx;
  s[1] = dvdy;
  s[2] = 0.5f * (dudy + dvdx);
}

// out0..2 = S_xx, S_yy, S_xy; with target != nullptr also nu = c_s^2 |S| (extract_les_samples_2d)
__global__ __launch_bounds__(256) void k_strain(const float* __restrict__ ux, const float* __restrict__ uy, int ny,
                                                int nx, float* __restrict__ feats, float* __restrict__ target,
                                                float cs2) {
  const int64_t n = int64_t(ny) * nx;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float s[3];
  strain_at(ux, uy, ny, nx, int(c / nx), int(c % nx), s);
  feats[c * 3 + 0] = s[0];
  feats[c * 3 + 1] = s[1];
  feats[c * 3 + 2] = s[2];
  if (target) target[c] = cs2 * sqrtf(2.0f * (s[0] * s[0] + s[1] * s[1] + 2.0f * s[2] * s[2]));
}

// mode 0: nu_t = max(model, 0) (predict_nu_t_2d); mode 1: tau_eff = max(tau + 3 nu_t, 0.501)
// (predict_tau_eff_2d, ai/inference.py:53-65)
template <int H>
__global__ __launch_bounds__(256) void k_nu_t(const float* __restrict__ p, Arch a, const float* __restrict__ ux,
                                              const float* __restrict__ uy, int ny, int nx, int mode, float tau,
                                              float* __restrict__ out) {
  const int64_t n = int64_t(ny) * nx;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float s[MAX_IN] = {};
  float s3[3];
  strain_at(ux, uy, ny, nx, int(c / nx), int(c % nx), s3);
  s[0] = s3[0];
  s[1] = s3[1];
  s[2] = s3[2];
  float nu = softplus(forward<H>(p, a, s, nullptr));
  nu = nu > 0.0f ? nu : 0.0f;
  if (mode == 0) {
    out[c] = nu;
  } else {
    const float te = tau + 3.0f * nu;
    out[c] = te > 0.501f ? te : 0.501f;
  }
}

// collide_ai_les_bgk (ai/inference.py:68-92): f - (f - f_eq(rho, u)) / tau_eff with tau_eff of the
// cell's strain; rho, ux, uy are the macroscopic fields of f (macroscopic, d2q9.py:61-74)
template <int H>
__global__ __launch_bounds__(256) void k_ai_les_bgk(const float* __restrict__ f, float* __restrict__ out,
                                                    const float* __restrict__ rho, const float* __restrict__ ux,
                                                    const float* __restrict__ uy, int ny, int nx,
                                                    const float* __restrict__ p, Arch a, float tau) {
  const int64_t n = int64_t(ny) * nx;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float s[MAX_IN] = {};
  float s3[3];
  strain_at(ux, uy, ny, nx, int(c / nx), int(c % nx), s3);
  s[0] = s3[0];
  s[1] = s3[1];
  s[2] = s3[2];
  float nu = softplus(forward<H>(p, a, s, nullptr));
  nu = nu > 0.0f ? nu : 0.0f;
  float te = tau + 3.0f * nu;
  te = te > 0.501f ? te : 0.501f;
  const float r = rho[c], u = ux[c], v = uy[c];
  const float usq = u * u + v * v;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const float cu = float(D2Q9::cx[q]) * u + float(D2Q9::cy[q]) * v;
    const float feq = float(D2Q9::w[q]) * r * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * usq);
    const float fq = f[q * n + c];
    out[q * n + c] = fq - (fq - feq) / te;
  }
}

// MSE gradient partials: partials[block][nparam + 2] = parameter gradients of the block's
// samples, then its loss share and its sum of |pred - target|
template <int H>
__global__ __launch_bounds__(256) void k_grad(const float* __restrict__ p, Arch a, const float* __restrict__ x,
                                              const float* __restrict__ t, uint64_t n, float inv_n,
                                              float* __restrict__ partials) {
  extern __shared__ float lds[];
  const int L = a.layers;
  // per sample: normalised input [MAX_IN], pre-activations [L][H], deltas [L][H], output delta,
  // loss share, |error|
  float* s_x = lds;
  float* s_z = s_x + BS * MAX_IN;
  float* s_d = s_z + BS * L * H;
  float* s_d3 = s_d + BS * L * H;
  float* s_loss = s_d3 + BS;
  float* s_abs = s_loss + BS;
  const int tid = threadIdx.x;
  const int64_t np = nparam(a);
  if (tid < BS) {
    const uint64_t s = uint64_t(blockIdx.x) * BS + tid;
    const bool ok = s < n;
    float xs[MAX_IN] = {};
    const float* mean = p + np;
    const float* sd = mean + a.in;
    for (int e = 0; e < a.in; ++e) xs[e] = ok ? x[s * a.in + e] : mean[e];
    float* zrow = s_z + tid * L * H;
    const float z3 = forward<H>(p, a, xs, zrow);
    for (int e = 0; e < MAX_IN; ++e) s_x[tid * MAX_IN + e] = e < a.in ? (xs[e] - mean[e]) / sd[e] : 0.0f;
    const float y = softplus(z3);
    const float res = ok ? y - t[s] : 0.0f;
    s_loss[tid] = res * res * inv_n;
    s_abs[tid] = ok ? fabsf(res) : 0.0f;
    const float bz = BETA * z3;
    const float d3 = 2.0f * res * inv_n * (bz > 20.0f ? 1.0f : 1.0f / (1.0f + expf(-bz)));
    s_d3[tid] = d3;
    // backward through the hidden layers
    const float* w_out = p + H * (a.in + 1) + int64_t(L - 1) * H * (H + 1);
    float g[H];
#pragma unroll
    for (int j = 0; j < H; ++j) g[j] = d3 * w_out[j] * dact(a.act, zrow[(L - 1) * H + j]);
    float* drow = s_d + tid * L * H;
    for (int l = L - 1;; --l) {
#pragma unroll
      for (int j = 0; j < H; ++j) drow[l * H + j] = g[j];
      if (l == 0) break;
      const float* w = p + H * (a.in + 1) + int64_t(l - 1) * H * (H + 1);  // W of layer l: [H][H]
      float gp[H];
#pragma unroll
      for (int i = 0; i < H; ++i) {
        float s2 = 0.0f;
#pragma unroll
        for (int j = 0; j < H; ++j) s2 = fmaf(g[j], w[j * H + i], s2);
        gp[i] = s2 * dact(a.act, zrow[(l - 1) * H + i]);
      }
#pragma unroll
      for (int i = 0; i < H; ++i) g[i] = gp[i];
    }
  }
  __syncthreads();
  float* out = partials + uint64_t(blockIdx.x) * (np + 2);
  const int64_t first_hidden = int64_t(H) * (a.in + 1);
  for (int64_t k = tid; k < np + 2; k += blockDim.x) {
    float acc = 0.0f;
    if (k == np) {
      for (int s = 0; s < BS; ++s) acc += s_loss[s];
    } else if (k == np + 1) {
      for (int s = 0; s < BS; ++s) acc += s_abs[s];
    } else if (k < first_hidden) {  // first layer: W [H][in], b [H]
      if (k < int64_t(H) * a.in) {
        const int j = int(k / a.in), e = int(k % a.in);
        for (int s = 0; s < BS; ++s) acc += s_d[s * L * H + j] * s_x[s * MAX_IN + e];
      } else {
        const int j = int(k - int64_t(H) * a.in);
        for (int s = 0; s < BS; ++s) acc += s_d[s * L * H + j];
      }
    } else if (k < first_hidden + int64_t(L - 1) * H * (H + 1)) {  // hidden layer l >= 1
      const int64_t r = k - first_hidden;
      const int l = int(r / (H * (H + 1))) + 1;
      const int64_t o = r % (H * (H + 1));
      if (o < int64_t(H) * H) {
        const int j = int(o / H), i = int(o % H);
        for (int s = 0; s < BS; ++s) {
          const float* zr = s_z + s * L * H + (l - 1) * H;
          acc += s_d[s * L * H + l * H + j] * act(a.act, zr[i]);
        }
      } else {
        const int j = int(o - int64_t(H) * H);
        for (int s = 0; s < BS; ++s) acc += s_d[s * L * H + l * H + j];
      }
    } else {  // output layer: w [H], b
      const int64_t o = k - first_hidden - int64_t(L - 1) * H * (H + 1);
      if (o < H) {
        for (int s = 0; s < BS; ++s) acc += s_d3[s] * act(a.act, s_z[s * L * H + (L - 1) * H + int(o)]);
      } else {
        for (int s = 0; s < BS; ++s) acc += s_d3[s];
      }
    }
    out[k] = acc;
  }
}

// Sum the partials in block order, clip the gradient's global 2-norm (clip_grad_norm_ rule:
// scale by min(1, clip / (|g| + 1e-6))), one Adam step (torch.optim.Adam rule); stats[0] = loss
// (MSE), stats[1] = sum |error|
constexpr int ADAM_THREADS = 1024;
__global__ __launch_bounds__(ADAM_THREADS) void k_adam(float* __restrict__ p, const float* __restrict__ partials,
                                                       int nblk, int64_t np, float* __restrict__ m,
                                                       float* __restrict__ v, float lr, float b1, float b2, float eps,
                                                       float bc1, float bc2, float clip, float* __restrict__ stats) {
  __shared__ float s_red[ADAM_THREADS];
  const int tid = threadIdx.x;
  float sq = 0.0f;
  for (int64_t k = tid; k < np; k += ADAM_THREADS) {
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g += partials[uint64_t(b) * (np + 2) + k];
    sq += g * g;
  }
  s_red[tid] = sq;
  __syncthreads();
  for (int off = ADAM_THREADS / 2; off > 0; off >>= 1) {
    if (tid < off) s_red[tid] += s_red[tid + off];
    __syncthreads();
  }
  const float norm = sqrtf(s_red[0]);
  const float scale = clip > 0.0f ? fminf(1.0f, clip / (norm + 1e-6f)) : 1.0f;
  if (tid < 2 && stats) {
    float a = 0.0f;
    for (int b = 0; b < nblk; ++b) a += partials[uint64_t(b) * (np + 2) + np + tid];
    stats[tid] = a;
  }
  if (lr == 0.0f) return;
  for (int64_t k = tid; k < np; k += ADAM_THREADS) {
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g += partials[uint64_t(b) * (np + 2) + k];
    g *= scale;
    const float mk = b1 * m[k] + (1.0f - b1) * g;
    const float vk = b2 * v[k] + (1.0f - b2) * g * g;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
  }
}

}  // namespace mlp2d
}  // namespace tlbm

// ===========================================================================
// C ABI
// ===========================================================================

using namespace tlbm;
using namespace tlbm::mlp2d;

namespace {

int c2fail(int code, const char* msg) { return tlbm::set_error(code, msg); }

int c2launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    return tlbm::set_error(TLBM_ERR_HIP, buf);
  }
  return TLBM_OK;
}

int check_arch(const tlbm_mlp2d_arch* arch, Arch* a) {
  if (!arch) return c2fail(TLBM_ERR_INVALID, "null architecture");
  a->in = arch->in_features;
  a->hidden = arch->hidden_features;
  a->layers = arch->hidden_layers;
  a->act = arch->activation;
  if (a->in < 1 || a->in > MAX_IN) return c2fail(TLBM_ERR_UNSUPPORTED, "in_features must be 1..8");
  if (a->hidden != 8 && a->hidden != 16 && a->hidden != 32 && a->hidden != 64)
    return c2fail(TLBM_ERR_UNSUPPORTED, "hidden_features must be 8, 16, 32 or 64");
  if (a->layers < 1 || a->layers > MAX_LAYERS) return c2fail(TLBM_ERR_UNSUPPORTED, "n_hidden_layers must be 1..4");
  if (a->act != TLBM_ACT_TANH && a->act != TLBM_ACT_RELU && a->act != TLBM_ACT_GELU)
    return c2fail(TLBM_ERR_INVALID, "unknown activation");
  return TLBM_OK;
}

// run fn.template operator()<H>() for the architecture's hidden width
template <class F>
int by_width(const Arch& a, F&& fn) {
  switch (a.hidden) {
    case 8: return fn(std::integral_constant<int, 8>{});
    case 16: return fn(std::integral_constant<int, 16>{});
    case 32: return fn(std::integral_constant<int, 32>{});
    default: return fn(std::integral_constant<int, 64>{});
  }
}

uint32_t blocks256(uint64_t n) { return uint32_t((n + 255) / 256); }

}  // namespace

extern "C" {

int64_t tlbm_mlp2d_nparam(const tlbm_mlp2d_arch* arch) {
  Arch a;
  if (check_arch(arch, &a) != TLBM_OK) return -1;
  return nparam(a);
}

int tlbm_mlp2d_forward(const tlbm_mlp2d_arch* arch, const float* params, const float* x, int64_t n, float* y,
                       tlbm_stream_t stream) {
  Arch a;
  int rc = check_arch(arch, &a);
  if (rc) return rc;
  if (n < 0) return c2fail(TLBM_ERR_SHAPE, "negative sample count");
  if (n == 0) return TLBM_OK;
  if (!params || !x || !y) return c2fail(TLBM_ERR_INVALID, "null pointer");
  if (uint64_t(n) >= (uint64_t(1) << 38)) return c2fail(TLBM_ERR_SHAPE, "too many samples");
  return by_width(a, [&](auto hw) {
    constexpr int H = decltype(hw)::value;
    hipLaunchKernelGGL(k_forward<H>, dim3(blocks256(uint64_t(n))), dim3(256), 0, (hipStream_t)stream, params, a, x,
                       uint64_t(n), y);
    return c2launch("k_forward");
  });
}

int tlbm_strain2d(const float* ux, const float* uy, int ny, int nx, float* feats, float* target, double c_s,
                  tlbm_stream_t stream) {
  if (ny < 1 || nx < 1) return c2fail(TLBM_ERR_SHAPE, "grid dims must be >= 1");
  if (!ux || !uy || !feats) return c2fail(TLBM_ERR_INVALID, "null pointer");
  const uint64_t n = uint64_t(ny) * uint64_t(nx);
  if (n >= (uint64_t(1) << 31)) return c2fail(TLBM_ERR_SHAPE, "grid exceeds 2^31 cells");
  hipLaunchKernelGGL(k_strain, dim3(blocks256(n)), dim3(256), 0, (hipStream_t)stream, ux, uy, ny, nx, feats, target,
                     float(c_s * c_s));
  return c2launch("k_strain");
}

int tlbm_mlp2d_nu_t(const tlbm_mlp2d_arch* arch, const float* params, const float* ux, const float* uy, int ny,
                    int nx, int mode, double tau, float* out, tlbm_stream_t stream) {
  Arch a;
  int rc = check_arch(arch, &a);
  if (rc) return rc;
  if (a.in != 3) return c2fail(TLBM_ERR_SHAPE, "the strain features need in_features = 3");
  if (ny < 1 || nx < 1) return c2fail(TLBM_ERR_SHAPE, "grid dims must be >= 1");
  if (!params || !ux || !uy || !out) return c2fail(TLBM_ERR_INVALID, "null pointer");
  if (mode != 0 && mode != 1) return c2fail(TLBM_ERR_INVALID, "mode must be 0 (nu_t) or 1 (tau_eff)");
  const uint64_t n = uint64_t(ny) * uint64_t(nx);
  if (n >= (uint64_t(1) << 31)) return c2fail(TLBM_ERR_SHAPE, "grid exceeds 2^31 cells");
  return by_width(a, [&](auto hw) {
    constexpr int H = decltype(hw)::value;
    hipLaunchKernelGGL(k_nu_t<H>, dim3(blocks256(n)), dim3(256), 0, (hipStream_t)stream, params, a, ux, uy, ny, nx,
                       mode, float(tau), out);
    return c2launch("k_nu_t");
  });
}

int tlbm_ai_les_bgk(const tlbm_mlp2d_arch* arch, const float* params, const float* f_in, float* f_out,
                    const float* rho, const float* ux, const float* uy, int ny, int nx, double tau,
                    tlbm_stream_t stream) {
  Arch a;
  int rc = check_arch(arch, &a);
  if (rc) return rc;
  if (a.in != 3) return c2fail(TLBM_ERR_SHAPE, "the strain features need in_features = 3");
  if (ny < 1 || nx < 1) return c2fail(TLBM_ERR_SHAPE, "grid dims must be >= 1");
  if (!params || !f_in || !f_out || !rho || !ux || !uy) return c2fail(TLBM_ERR_INVALID, "null pointer");
  if (f_in == f_out) return c2fail(TLBM_ERR_ALIAS, "f_in and f_out alias");
  if (!(tau > 0.5)) return c2fail(TLBM_ERR_INVALID, "tau must be > 0.5");
  const uint64_t n = uint64_t(ny) * uint64_t(nx);
  if (n >= (uint64_t(1) << 31)) return c2fail(TLBM_ERR_SHAPE, "grid exceeds 2^31 cells");
  return by_width(a, [&](auto hw) {
    constexpr int H = decltype(hw)::value;
    hipLaunchKernelGGL(k_ai_les_bgk<H>, dim3(blocks256(n)), dim3(256), 0, (hipStream_t)stream, f_in, f_out, rho, ux,
                       uy, ny, nx, params, a, float(tau));
    return c2launch("k_ai_les_bgk");
  });
}

int tlbm_mlp2d_grad(const tlbm_mlp2d_arch* arch, const float* params, const float* x, const float* t, int64_t n,
                    float* partials, tlbm_stream_t stream) {
  Arch a;
  int rc = check_arch(arch, &a);
  if (rc) return rc;
  if (n <= 0) return c2fail(TLBM_ERR_SHAPE, "need at least one sample");
  if (!params || !x || !t || !partials) return c2fail(TLBM_ERR_INVALID, "null pointer");
  const uint64_t nb = (uint64_t(n) + BS - 1) / BS;
  if (nb >= (uint64_t(1) << 31)) return c2fail(TLBM_ERR_SHAPE, "batch too large");
  const size_t lds = sizeof(float) * (size_t(BS) * MAX_IN + 2 * size_t(BS) * a.layers * a.hidden + 3 * BS);
  if (lds > 65536)
    return c2fail(TLBM_ERR_UNSUPPORTED, "GPU training keeps 64 samples' activations in 64 KB of LDS: "
                                        "hidden_features * n_hidden_layers must be at most 120");
  return by_width(a, [&](auto hw) {
    constexpr int H = decltype(hw)::value;
    hipLaunchKernelGGL(k_grad<H>, dim3(uint32_t(nb)), dim3(256), lds, (hipStream_t)stream, params, a, x, t,
                       uint64_t(n), 1.0f / float(n), partials);
    return c2launch("k_grad");
  });
}

int tlbm_mlp2d_adam(const tlbm_mlp2d_arch* arch, float* params, const float* partials, int nblocks, float* m,
                    float* v, int step, float lr, float beta1, float beta2, float eps, float clip_norm, float* stats,
                    tlbm_stream_t stream) {
  Arch a;
  int rc = check_arch(arch, &a);
  if (rc) return rc;
  if (nblocks <= 0 || step < 1) return c2fail(TLBM_ERR_INVALID, "need nblocks > 0 and step >= 1");
  if (!params || !partials || (lr != 0.0f && (!m || !v))) return c2fail(TLBM_ERR_INVALID, "null pointer");
  const float bc1 = 1.0f - powf(beta1, float(step));
  const float bc2 = 1.0f - powf(beta2, float(step));
  hipLaunchKernelGGL(k_adam, dim3(1), dim3(ADAM_THREADS), 0, (hipStream_t)stream, params, partials, nblocks,
                     nparam(a), m, v, lr, beta1, beta2, eps, bc1, bc2, clip_norm, stats);
  return c2launch("k_adam");
}

}  // extern "C"
