// Population storage formats.  Arithmetic is float for the 32/16-bit
// storages and double for binary64.  The deviation storages keep
// g_i = f_i - w_i, so the 16-bit quantisation acts on the small
// non-equilibrium part instead of the O(w_i) population.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.hpp"

namespace tlbm {

enum StoreKind { ST_F32 = 0, ST_F64 = 1, ST_F16 = 2, ST_BF16 = 3, ST_F16_DEV = 4, ST_BF16_DEV = 5 };

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// round to nearest even (the hardware conversion, v_cvt_pk_bf16_f32), NaN kept quiet
__device__ __forceinline__ uint16_t f32_to_bf16(float v) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v)); }


template <int S>
struct Store;

template <>
struct Store<ST_F32> {
  using T = float;
  using C = float;
  template <class L>
  __device__ static __forceinline__ C get(T v, int) { return v; }
  template <class L>
  __device__ static __forceinline__ T put(C v, int) { return v; }
  static constexpr bool kDev = false;
};

template <>
struct Store<ST_F64> {
  using T = double;
  using C = double;
  template <class L>
  __device__ static __forceinline__ C get(T v, int) { return v; }
  template <class L>
  __device__ static __forceinline__ T put(C v, int) { return v; }
  static constexpr bool kDev = false;
};

template <>
struct Store<ST_F16> {
  using T = _Float16;
  using C = float;
  template <class L>
  __device__ static __forceinline__ C get(T v, int) { return float(v); }
  template <class L>
  __device__ static __forceinline__ T put(C v, int) { return _Float16(v); }
  static constexpr bool kDev = false;
};

template <>
struct Store<ST_BF16> {
  using T = uint16_t;
  using C = float;
  template <class L>
  __device__ static __forceinline__ C get(T v, int) { return bf16_to_f32(v); }
  template <class L>
  __device__ static __forceinline__ T put(C v, int) { return f32_to_bf16(v); }
  static constexpr bool kDev = false;
};

template <>
struct Store<ST_F16_DEV> {
  using T = _Float16;
  using C = float;
  template <class L>
  __device__ static __forceinline__ C get(T v, int q) { return float(v) + float(L::w[q]); }
  template <class L>
  __device__ static __forceinline__ T put(C v, int q) { return _Float16(v - float(L::w[q])); }
  static constexpr bool kDev = true;
  // the stored deviation itself, for a step that collides in deviations
  __device__ static __forceinline__ C get_dev(T v) { return float(v); }
  // v already holds f - w (the BGK relaxation's deviation output).  The empty asm
  // keeps the conversion a v_cvt_f16_f32 of the fp32 value: the relaxation's last
  // fma and the conversion would otherwise fuse into one v_fma_mixlo_f16 (a single
  // rounding) in some kernels and not in others, and the kernels would disagree.
  __device__ static __forceinline__ T put_dev(C v) {
    asm("" : "+v"(v));
    return _Float16(v);
  }
};

template <>
struct Store<ST_BF16_DEV> {
  using T = uint16_t;
  using C = float;
  template <class L>
  __device__ static __forceinline__ C get(T v, int q) { return bf16_to_f32(v) + float(L::w[q]); }
  template <class L>
  __device__ static __forceinline__ T put(C v, int q) { return f32_to_bf16(v - float(L::w[q])); }
  static constexpr bool kDev = true;
  __device__ static __forceinline__ C get_dev(T v) { return bf16_to_f32(v); }
  __device__ static __forceinline__ T put_dev(C v) { return f32_to_bf16(v); }
};

}  // namespace tlbm
