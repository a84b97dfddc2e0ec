// libtlbm_hip: the per-step lattice update of TensorLBM for MI355X (gfx950).
//
// One kernel template (k_lbm) covers the fused pull-stream + boundary +
// bounce-back + collide step, the stream-only op and the collide-only op for
// D2Q9 / D3Q19 / D3Q27 and all storage formats.
//
// Work mapping: a workgroup of 256 lanes covers a (BX x BY) tile of one
// z-plane, x fastest, one cell per lane.  Each lane gathers its Q incoming
// populations (coalesced along x; the +-1 x shift costs one partial cache line
// per row, absorbed by L2 so HBM traffic stays at 2 Q itemsize per cell),
// collides in registers and streams Q values out with non-temporal stores.
//
// Addressing: populations may use any (q, z, y) strides with unit x stride.
// For lane q the plane base  f + q*sq + z_src*sz  is uniform over the
// workgroup (64-bit, scalar registers) and only  y_src*sy + x_src  is a
// per-lane 32-bit offset, so every load/store uses the scalar-base +
// 32-bit-offset form.  The solver keeps its buffers in a [z][q][y][x] layout
// with a padded q stride (fewer concurrently open DRAM rows than plain SoA);
// the reference's contiguous [q][z][y][x] tensors are served by the same code.
//
// Full-way bounce-back at a solid cell is a register lane swap
// f[q] <-> f[opp(q)] of the pulled values (pull[opp q](x) = b[opp q][x + c_q]),
// so walls cost no extra loads.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <type_traits>

#include "../../include/tlbm_hip.h"
#include "collide.hpp"
#include "mlp.hpp"
#include "store.hpp"

namespace tlbm {

// FEAT_BC: solid mask and boundaries; FEAT_FORCE: Ladd force sample; FEAT_MLP:
// learned closure (whole waves stay alive for the matrix-core evaluation);
// FEAT_MLP_BF16 with FEAT_MLP: its bf16-operand form.
// Bits combine; the one-step kernel builds FEAT_ALL, the two-step kernel FEAT_BC.
enum Feat {
  FEAT_NONE = 0,
  FEAT_BC = 1,
  FEAT_MLP = 2,
  FEAT_FORCE = 4,
  FEAT_MLP_BF16 = 8,
  FEAT_ALL = FEAT_BC | FEAT_FORCE,
  FEAT_MLP_ALL = FEAT_MLP | FEAT_ALL
};

// boundary kind known at compile time (a TLBM_BC_* value) or read from the launch
constexpr int BCK_RUNTIME = -1;

template <typename C>
struct KArgs {
  const void* fin;
  void* fout;
  const int8_t* solid;
  const float* tau_eff;
  const void* mlp;
  double* force;
  int64_t isq, isz, isy;  // input strides (elements; x stride 1)
  int64_t osq, osz, osy;  // output strides
  int nx, ny, nz, tiles_x, tiles_y;
  int tau_mode, eq_floor, bc_kind, bc_faces;
  C omega, tau, cs_delta_sq, tau_cap;
  C rate[8];
  C bc_u[3];
};

// Population buffers are device (global) memory.  Saying so in the pointer type
// makes the compiler emit global_load / global_store; through a generic pointer
// it emits FLAT operations, which count in both the vector-memory and the LDS
// counters, so every wait for LDS (the workgroup barrier's) would also wait for
// the outstanding stores to reach memory.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;

// value at a uniform base plus a per-lane 32-bit element offset
template <typename T>
__device__ __forceinline__ T ld_off(const T* base, uint32_t off) {
  return *(global_ptr<const T>)(reinterpret_cast<const char*>(base) + off * uint32_t(sizeof(T)));
}

// A pointer the caller knows to be uniform over the wave, moved to scalar registers
// so that the memory op uses the scalar-base + 32-bit-offset form (otherwise the
// compiler may build a 64-bit per-lane address with quarter-rate multiplies).
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
  return reinterpret_cast<T*>((uint64_t(hi) << 32) | lo);
}

// non-temporal store to a wave-uniform base plus a per-lane 32-bit element offset
template <typename T>
__device__ __forceinline__ void st_nt(T* base, uint32_t off, T v) {
  __builtin_nontemporal_store(v, (global_ptr<T>)(reinterpret_cast<char*>(uniform_ptr(base)) + off * uint32_t(sizeof(T))));
}

// the same address through L2 (write-back): partial cache lines written by two
// workgroups merge there before they reach HBM
template <typename T>
__device__ __forceinline__ void st_l2(T* base, uint32_t off, T v) {
  *(global_ptr<T>)(reinterpret_cast<char*>(uniform_ptr(base)) + off * uint32_t(sizeof(T))) = v;
}

__device__ __forceinline__ int wrap_dn(int v, int n) { return v == 0 ? n - 1 : v - 1; }
__device__ __forceinline__ int wrap_up(int v, int n) { return v == n - 1 ? 0 : v + 1; }

// source coordinate of lane velocity component c for a pull: x - c with wrap
__device__ __forceinline__ int src_of(int c, int dn, int mid, int up) { return c > 0 ? dn : (c < 0 ? up : mid); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// pulled population q of cell (x, y, z) (periodic wrap), used off the hot loop;
// `Uniform` as for gather (z is wave-uniform wherever it is set)
template <class ST, class L, typename C, bool Uniform = false>
__device__ __forceinline__ typename ST::C pull_one(const typename ST::T* fin, const KArgs<C>& a, int q, int x, int y,
                                                   int z) {
  int sx = x - L::cx[q], sy = y - L::cy[q], sz = z - L::cz[q];
  sx = sx < 0 ? sx + a.nx : (sx >= a.nx ? sx - a.nx : sx);
  sy = sy < 0 ? sy + a.ny : (sy >= a.ny ? sy - a.ny : sy);
  sz = sz < 0 ? sz + a.nz : (sz >= a.nz ? sz - a.nz : sz);
  const typename ST::T* base = fin + int64_t(q) * a.isq + int64_t(sz) * a.isz;
  if constexpr (Uniform) base = uniform_ptr(base);
  return ST::template get<L>(ld_off(base, uint32_t(sy) * uint32_t(a.isy) + uint32_t(sx)), q);
}

extern __shared__ __align__(16) unsigned char dyn_lds[];

#ifndef TLBM_MLP_WAVES
#define TLBM_MLP_WAVES 4
#endif

// 1/x by the hardware reciprocal (1 ulp) in fp32; exact division in fp64
__device__ __forceinline__ float approx_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double approx_rcp(double x) { return 1.0 / x; }
// square root by the hardware instruction (1 ulp) in fp32; exact in fp64
__device__ __forceinline__ float approx_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double approx_sqrt(double x) { return sqrt(x); }

// Equilibrium of the boundary velocity u (a launch constant) computed where it
// is used.  The empty asm makes the compiler treat the velocity as produced at
// this point: otherwise it hoists the Q equilibrium values out of the z loop of
// the two-step kernel and keeps them in registers for its whole run.
template <class L, typename C>
__device__ __forceinline__ void equilibrium_here(C rho, const C (&u)[3], C (&f)[L::Q]) {
  C ux = u[0], uy = u[1], uz = u[2];
  asm volatile("" : "+v"(ux), "+v"(uy), "+v"(uz));
  equilibrium<L>(rho, ux, uy, uz, f);
}

// How update_cell sees the pulled states of the neighbouring column that the
// channel inlet (x = 1) and the zero-gradient outlets (x = nx-2) read:
//   BCN_PULL : through the callback pulled(q, x, y, z) (one-step kernel, step
//              two of the two-step kernel)
//   BCN_IN_F : on those edge cells (`nb_edge`) f already holds the neighbouring
//              column's pulled state, gathered there instead of at the cell
//              (step one of the two-step kernel: no extra loads, no extra registers)
enum BcNeighbour { BCN_PULL = 0, BCN_IN_F = 1 };

// Ladd sample of cell (x, y, z)'s own pulled populations read from the input
// lattice, for the rare solid edge cell whose f holds its neighbour's state
// (rolled loop: kept off the register budget of the common path)
template <class L, int S>
__device__ __forceinline__ void own_force_sample(const KArgs<typename Store<S>::C>& a, int x, int y, int z,
                                                 double& fx, double& fy, double& fz) {
  const typename Store<S>::T* fin = static_cast<const typename Store<S>::T*>(a.fin);
#pragma unroll 1
  for (int q = 1; q < L::Q; ++q) {
    const double v = double(pull_one<Store<S>, L>(fin, a, q, x, y, z));
    if (L::cx[q] != 0) fx += double(L::cx[q]) * v;
    if (L::cy[q] != 0) fy += double(L::cy[q]) * v;
    if (L::cz[q] != 0) fz += double(L::cz[q]) * v;
  }
}

// Post-stream work of one cell, shared by the one-step and the two-step kernels:
// Ladd force sample, boundary overrides, full-way bounce-back, collision.
//
// `eq_unit(e)` fills e with f_eq(rho = 1, bc_u), from which the inlet takes
// rho_1 * e (f_eq is linear in rho) and the far-field faces take e itself.
//
// The learned closure reads its operand blob from `mlp_blob` when the kernel
// staged it (an LDS pointer), else from the global blob a.mlp.
// `before_collision()` runs after the boundary loads and bounce-back, before the
// collision (the closure kernel issues its next plane's loads there).
// The BGK collision of the deviation storage forms hands back f - w (the
// relaxation absorbs the shift), so their stores convert only; every other
// collision, and streaming alone, hands back f.
template <int S, int COLL>
constexpr bool dev_out() {
  return Store<S>::kDev && COLL == TLBM_COLL_BGK;
}
template <class L, int S, int COLL>
__device__ __forceinline__ typename Store<S>::T put_out(typename Store<S>::C v, int q) {
  if constexpr (dev_out<S, COLL>()) return Store<S>::put_dev(v);
  else return Store<S>::template put<L>(v, q);
}

struct NoHook {
  __device__ void operator()() const {}
};
// kDevIn: f (and what `pulled` returns) holds the deviations f - w_q of a 16-bit
// deviation store through to the collision; the closure march's BGK step only.
template <class L, int S, int COLL, bool STREAM, int FEAT, int BCN, int BCK, bool kDevIn = false, class Pulled,
          class EqUnit, class Blob = const mlp::Packed*, class Hook = NoHook>
__device__ __forceinline__ void update_cell(typename Store<S>::C (&f)[L::Q], const KArgs<typename Store<S>::C>& a,
                                            int xx, int yy, int z, uint32_t cell, bool solid, bool nb_edge,
                                            double& fx, double& fy, double& fz, const Pulled& pulled,
                                            const EqUnit& eq_unit, Blob mlp_blob = nullptr,
                                            const Hook& before_collision = Hook()) {
  using C = typename Store<S>::C;
  constexpr int Q = L::Q;
  constexpr bool kAll = (FEAT & FEAT_BC) != 0;
  constexpr bool kForce = (FEAT & FEAT_FORCE) != 0;
  constexpr bool kMlp = (FEAT & FEAT_MLP) != 0;
  static_assert(!kDevIn || (dev_out<S, COLL>() && kMlp), "deviations through the closure's BGK step only");
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  // a compile-time kind folds the other boundaries away
  const int bck = BCK == BCK_RUNTIME ? a.bc_kind : BCK;
  if constexpr (kAll) {
    if (kForce && a.force != nullptr && solid) {
      // Ladd sample: post-stream, pre-bounce-back (obstacles.py:91-130)
      if (BCN == BCN_IN_F && nb_edge) {
        own_force_sample<L, S>(a, xx, yy, z, fx, fy, fz);
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (L::cx[q] != 0) fx += double(L::cx[q]) * double(f[q]);
          if (L::cy[q] != 0) fy += double(L::cy[q]) * double(f[q]);
          if (L::cz[q] != 0) fz += double(L::cz[q]) * double(f[q]);
        }
      }
    }
    if (STREAM && bck == TLBM_BC_CHANNEL_X) {
      if (xx == 0) {
        // equilibrium inlet, density of the pulled state at x = 1
        C rho1 = C(0);
        if constexpr (BCN == BCN_PULL) {
#pragma unroll
          for (int q = 0; q < Q; ++q) rho1 += pulled(q, 1, yy, z);
        } else {
#pragma unroll
          for (int q = 0; q < Q; ++q) rho1 += f[q];
        }
        if constexpr (kDevIn) {
          // rho1 is the neighbour's rho - 1 here: rho_1 e_q - w_q = e_q rho1 + (e_q - w_q)
          C e[Q], ed[Q];
          eq_unit(e);
          eq_unit.dev(ed);
#pragma unroll
          for (int q = 0; q < Q; ++q) f[q] = fma(e[q], rho1, ed[q]);
        } else {
          eq_unit(f);
#pragma unroll
          for (int q = 0; q < Q; ++q) f[q] *= rho1;
        }
      } else if (BCN == BCN_PULL && xx == nx - 1) {
        // zero-gradient outlet: copy the pulled state of x = nx-2
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = pulled(q, nx - 2, yy, z);
      }
    } else if (STREAM && bck == TLBM_BC_LID_Y && yy == ny - 1 && !solid) {
      // Zou-He on the fluid cells of the lid plane; solid cells there keep full-way bounce-back
      if constexpr (Q == 19) {
        if constexpr (kDevIn) {
#pragma unroll
          for (int q = 0; q < Q; ++q) f[q] += C(L::w[q]);
        }
        zou_he_lid(f, a.bc_u[0]);
        if constexpr (kDevIn) {
#pragma unroll
          for (int q = 0; q < Q; ++q) f[q] -= C(L::w[q]);
        }
      }
    } else if (STREAM && bck == TLBM_BC_FAR_FIELD) {
      // far_field_bc_3d (boundaries3d.py:446-537): free-stream equilibrium on the
      // inlet and the selected lateral faces (lateral faces win on edges),
      // zero-gradient outlet
      const int fm = a.bc_faces;
      const bool lateral = (yy == 0 && (fm & TLBM_FACE_YM)) || (yy == ny - 1 && (fm & TLBM_FACE_YP)) ||
                           (z == 0 && (fm & TLBM_FACE_ZM)) || (z == nz - 1 && (fm & TLBM_FACE_ZP));
      if (lateral || (xx == 0 && !(fm & TLBM_FACE_XM_PERIODIC))) {
        if constexpr (kDevIn) eq_unit.dev(f);
        else eq_unit(f);
      } else if (BCN == BCN_PULL && xx == nx - 1 && !(fm & TLBM_FACE_XP_PERIODIC)) {
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = pulled(q, nx - 2, yy, z);
      }
    }
    // full-way bounce-back; the wave vote skips the swaps in the waves that hold
    // no solid cell, i.e. almost all of them.  As per-lane selects the swapped
    // values stay in the registers of the unswapped ones, so the waves that skip
    // pay no copies at the join (a lane-masked exchange left 18 register copies
    // on every wave's path of the closure step); the force kernels and the plain
    // 16-bit deviation steps measured faster with the exchange.
    if (__any(solid)) {
      if constexpr (kForce || (Store<S>::kDev && !kMlp)) {
        if (solid) {
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            if (q < L::opp[q]) {
              const C t = f[q];
              f[q] = f[L::opp[q]];
              f[L::opp[q]] = t;
            }
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (q < L::opp[q]) {
            const C a = f[q], b = f[L::opp[q]];
            f[q] = solid ? b : a;
            f[L::opp[q]] = solid ? a : b;
          }
        }
      }
    }
  }
  before_collision();

  if constexpr (COLL != TLBM_COLL_NONE) {
    const Macro<C> mc = macro<L, C, kDevIn>(f);
    const C rho_eq = a.eq_floor ? mc.rho_safe : mc.rho;
    C omega = a.omega;
    const bool need_pi = (COLL == TLBM_COLL_CUMULANT && Q == 19) || a.tau_mode == TLBM_TAU_SMAG ||
                         a.tau_mode == TLBM_TAU_SMAG_RAW;
    if constexpr (COLL == TLBM_COLL_BGK && kMlp) {
      // learned closure: strain from the non-equilibrium stress with the
      // Smagorinsky tau as reference time scale, S = -Pi / (2 rho cs^2 tau_ref)
      C pi[6];
      neq_stress_moments<L, C, kDevIn>(f, rho_eq, mc.ux, mc.uy, mc.uz, pi);
      // features feed an fp16 network: the clamped Smagorinsky time scale (smagorinsky_tau)
      // and the strain scale take hardware reciprocals and square roots (1 ulp)
      const C pn = approx_sqrt(pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
                               C(2) * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]));
      C disc = fma(a.tau, a.tau, C(18) * a.cs_delta_sq * pn * approx_rcp(mc.rho_safe));
      disc = disc > C(0) ? disc : C(0);
      C te_ref = C(0.5) * (a.tau + approx_sqrt(disc));
      te_ref = te_ref < C(0.5001) ? C(0.5001) : te_ref;
      te_ref = te_ref > a.tau_cap ? a.tau_cap : te_ref;
      const C sc = C(-1.5) * approx_rcp(te_ref * mc.rho_safe);
      float xs[7];
#pragma unroll
      for (int k = 0; k < 6; ++k) xs[k] = float(sc * pi[k]);
      xs[6] = __builtin_amdgcn_sqrtf(2.0f * (xs[0] * xs[0] + xs[1] * xs[1] + xs[2] * xs[2] +
                                             2.0f * (xs[3] * xs[3] + xs[4] * xs[4] + xs[5] * xs[5])));
      // operands from the workgroup's LDS copy of the blob where the kernel staged one,
      // else straight from the (cache-resident, 3.5 KB) global blob
      float nu;
      if constexpr (std::is_same_v<Blob, mlp::LdsPacked>) {
        nu = (FEAT & FEAT_MLP_BF16) != 0 ? mlp::nu_t_bf16(mlp_blob, xs) : mlp::nu_t(mlp_blob, xs);
      } else {
        const mlp::Packed* blob = static_cast<const mlp::Packed*>(a.mlp);
        nu = (FEAT & FEAT_MLP_BF16) != 0 ? mlp::nu_t_bf16(blob, xs) : mlp::nu_t(blob, xs);
      }
      // tau_eff = tau + 3 nu_t, floored at 0.5 + 1e-3 (ai/inference.py:56-65)
      C te = a.tau + C(3) * C(nu);
      te = te < C(0.501) ? C(0.501) : te;
      bgk_relax<L, C, dev_out<S, COLL>(), kDevIn>(f, rho_eq, mc.ux, mc.uy, mc.uz, approx_rcp(te));
    } else if constexpr (COLL == TLBM_COLL_BGK) {
      if (need_pi) {
        C feq[Q], fneq[Q], pi[6];
        equilibrium<L>(rho_eq, mc.ux, mc.uy, mc.uz, feq);
#pragma unroll
        for (int q = 0; q < Q; ++q) fneq[q] = f[q] - feq[q];
        neq_stress<L>(fneq, pi);
        omega = a.tau_mode == TLBM_TAU_SMAG
                    ? smagorinsky_omega(pi_norm(pi), mc.rho_safe, a.tau, a.cs_delta_sq, a.tau_cap)
                    : smagorinsky_raw_omega(pi_norm(pi), mc.rho_safe, a.tau, a.cs_delta_sq);
      } else if (a.tau_mode == TLBM_TAU_FIELD) {
        omega = C(1) / C(a.tau_eff[cell]);
      }
      if constexpr (!STREAM) {
        if (a.tau_mode == TLBM_TAU_FIELD64) omega = C(1) / C(reinterpret_cast<const double*>(a.tau_eff)[cell]);
      }
      bgk_relax<L, C, dev_out<S, COLL>()>(f, rho_eq, mc.ux, mc.uy, mc.uz, omega);
    } else {
      // the Geier collision takes the equilibrium's moments in closed form and the
      // LES stress from the moments of f, so it never forms f_eq
      constexpr bool kFeq = COLL != TLBM_COLL_GEIER;
      C feq[Q];
      if constexpr (kFeq) equilibrium<L>(rho_eq, mc.ux, mc.uy, mc.uz, feq);
      C pi[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
      if (need_pi) {
        if constexpr (kFeq) {
          C fneq[Q];
#pragma unroll
          for (int q = 0; q < Q; ++q) fneq[q] = f[q] - feq[q];
          neq_stress<L>(fneq, pi);
        } else {
          neq_stress_moments<L>(f, rho_eq, mc.ux, mc.uy, mc.uz, pi);
        }
      }
      if (a.tau_mode == TLBM_TAU_SMAG) {
        omega = smagorinsky_omega(pi_norm(pi), mc.rho_safe, a.tau, a.cs_delta_sq, a.tau_cap);
      } else if (a.tau_mode == TLBM_TAU_SMAG_RAW) {
        omega = smagorinsky_raw_omega(pi_norm(pi), mc.rho_safe, a.tau, a.cs_delta_sq);
      } else if (a.tau_mode == TLBM_TAU_FIELD) {
        omega = C(1) / C(a.tau_eff[cell]);
      }
      if constexpr (!STREAM) {
        if (a.tau_mode == TLBM_TAU_FIELD64) omega = C(1) / C(reinterpret_cast<const double*>(a.tau_eff)[cell]);
      }
      if constexpr (COLL == TLBM_COLL_MRT) {
        mrt<L>(f, feq, a.rate, omega);
      } else if constexpr (COLL == TLBM_COLL_CM) {
        cascaded<L>(f, feq, mc, omega, a.rate);
      } else if constexpr (COLL == TLBM_COLL_CUMULANT) {
        if constexpr (Q == 19) cumulant19<L>(f, feq, pi, omega, a.rate);
        if constexpr (Q == 27) legacy27<C>(f, feq, mc, omega, a.rate);
      } else if constexpr (COLL == TLBM_COLL_GEIER) {
        if constexpr (Q == 27) geier27<C>(f, rho_eq, mc, omega, a.rate);
      }
    }
  }

}

// Q pulled populations of cell (x, y, z) from the input lattice (periodic wrap)
template <class L, int S>
__device__ __forceinline__ void gather(typename Store<S>::C (&f)[L::Q], const typename Store<S>::T* fin,
                                       const KArgs<typename Store<S>::C>& a, int x, int y, int z) {
  using ST = Store<S>;
  const int xd = wrap_dn(x, a.nx), xu = wrap_up(x, a.nx);
  const int yd = wrap_dn(y, a.ny), yu = wrap_up(y, a.ny);
  const int zd = wrap_dn(z, a.nz), zu = wrap_up(z, a.nz);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int sx = src_of(L::cx[q], xd, x, xu);
    const int sy = src_of(L::cy[q], yd, y, yu);
    const int sz = src_of(L::cz[q], zd, z, zu);
    const typename ST::T* base = fin + int64_t(q) * a.isq + int64_t(sz) * a.isz;
    f[q] = ST::template get<L>(ld_off(base, uint32_t(sy) * uint32_t(a.isy) + uint32_t(sx)), q);
  }
}

// The same loads as gather, left in the storage format (a prefetch converts them
// where they are used, so nothing waits for them at the load), each plane base
// moved to scalar registers: in a loop over z the compiler otherwise
// strength-reduces the Q addresses into 64-bit per-lane pointers held live
// across iterations.
template <class L, int S>
__device__ __forceinline__ void gather_raw(typename Store<S>::T (&r)[L::Q], const typename Store<S>::T* fin,
                                           const KArgs<typename Store<S>::C>& a, int x, int y, int z) {
  const int xd = wrap_dn(x, a.nx), xu = wrap_up(x, a.nx);
  const int yd = wrap_dn(y, a.ny), yu = wrap_up(y, a.ny);
  const int zd = wrap_dn(z, a.nz), zu = wrap_up(z, a.nz);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int sx = src_of(L::cx[q], xd, x, xu);
    const int sy = src_of(L::cy[q], yd, y, yu);
    const int sz = src_of(L::cz[q], zd, z, zu);
    const typename Store<S>::T* base = uniform_ptr(fin + int64_t(q) * a.isq + int64_t(sz) * a.isz);
    r[q] = ld_off(base, uint32_t(sy) * uint32_t(a.isy) + uint32_t(sx));
  }
}

// one FP64 atomic per wave that holds solid cells
__device__ __forceinline__ void force_flush(double* force, double fx, double fy, double fz, bool any) {
  if (__any(any)) {
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0) {
      atomicAdd(force + 0, 2.0 * fx);
      atomicAdd(force + 1, 2.0 * fy);
      atomicAdd(force + 2, 2.0 * fz);
    }
  }
}

// Whether the lane at x takes, before the update, the pulled state of the cell
// its boundary reads (x = 1 for the channel inlet, x = nx - 2 for a copy outlet),
// so that `pulled` can hand back its own f: only lanes whose boundary reads it.
// The far-field inlet and lateral faces set an equilibrium and read nothing; a
// periodic x face keeps its own state.
__device__ __forceinline__ bool take_neighbour(int bck, int faces, int x, int nx, bool out_shfl) {
  if (bck == TLBM_BC_CHANNEL_X) return x == 0 || (out_shfl && x == nx - 1);
  if (bck == TLBM_BC_FAR_FIELD) return out_shfl && x == nx - 1 && !(faces & TLBM_FACE_XP_PERIODIC);
  return false;
}

template <class L, int S, int COLL, bool STREAM, int FEAT, int BCK>
__device__ __forceinline__ void lbm_body(const KArgs<typename Store<S>::C>& a) {
  using ST = Store<S>;
  using T = typename ST::T;
  using C = typename ST::C;
  constexpr int Q = L::Q;
  constexpr bool kAll = (FEAT & FEAT_BC) != 0;

  const uint32_t tile = blockIdx.x;
  const uint32_t txi = tile % uint32_t(a.tiles_x);
  const uint32_t rest = tile / uint32_t(a.tiles_x);
  const uint32_t tyi = rest % uint32_t(a.tiles_y);
  const int z = int(rest / uint32_t(a.tiles_y));
  const int x = int(txi * blockDim.x + threadIdx.x);
  const int y = int(tyi * blockDim.y + threadIdx.y);
  const bool active = x < a.nx && y < a.ny;
  if (FEAT == FEAT_NONE && !active) return;
  // inactive lanes (force / closure paths) work on cell (0, 0, z): in bounds, never stored
  const int xx = active ? x : 0;
  const int yy = active ? y : 0;
  const uint32_t cell = (uint32_t(z) * uint32_t(a.ny) + uint32_t(yy)) * uint32_t(a.nx) + uint32_t(xx);
  const T* fin = static_cast<const T*>(a.fin);

  C f[Q];
  if constexpr (STREAM) {
    gather<L, S>(f, fin, a, xx, yy, z);
  } else {
    const uint32_t off = uint32_t(yy) * uint32_t(a.isy) + uint32_t(xx);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = ST::template get<L>(ld_off(fin + int64_t(q) * a.isq + int64_t(z) * a.isz, off), q);
  }
  // (issued after the gather: measured faster here than ahead of it, unlike the
  // two-step kernel's prefetch)
  int mask = 0;
  if constexpr (kAll) {
    if (a.solid != nullptr && active) mask = a.solid[cell];
  }
  const bool solid = mask != 0;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  // inlet / outlet neighbours by lane permute, as in lbm_mlp_body (inactive lanes
  // take part in the permute; what they compute is never stored)
  constexpr bool kForceFeat = (FEAT & FEAT_FORCE) != 0;
  const bool out_shfl = !kForceFeat && ((a.nx - 1) & 63) != 0;
  if constexpr (kAll && !kForceFeat && STREAM) {
    const int bck = BCK == BCK_RUNTIME ? a.bc_kind : BCK;
    const bool take = take_neighbour(bck, a.bc_faces, xx, a.nx, out_shfl);
    if (__any(take)) {
      const int src = xx == 0 ? 1 : ((a.nx - 2) & 63);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const C g = __shfl(f[q], src, 64);
        f[q] = take ? g : f[q];
      }
    }
  }
  auto pulled = [&](int q, int px, int py, int pz) -> C {
    if (!kForceFeat && (px == 1 || out_shfl)) return f[q];
    return pull_one<ST, L>(fin, a, q, px, py, pz);
  };
  auto eq_unit = [&](C (&e)[Q]) { equilibrium_here<L>(C(1), a.bc_u, e); };
  update_cell<L, S, COLL, STREAM, FEAT, BCN_PULL, BCK>(f, a, xx, yy, z, cell, solid, false, fx, fy, fz, pulled,
                                                        eq_unit);
  if (active) {
    T* fout = static_cast<T*>(a.fout);
    const uint32_t off = uint32_t(yy) * uint32_t(a.osy) + uint32_t(xx);
#pragma unroll
    for (int q = 0; q < Q; ++q) st_nt(fout + int64_t(q) * a.osq + int64_t(z) * a.osz, off, put_out<L, S, COLL>(f[q], q));
  }
  if constexpr ((FEAT & FEAT_FORCE) != 0) {
    if (a.force != nullptr) force_flush(a.force, fx, fy, fz, solid);
  }
}

// Workgroup w of a 1-D grid of n runs on XCD w % 8 (round-robin dispatch over the
// 8 dies, each with its own L2).  Give XCD c the contiguous tile range c*n/8 ..
// so that tiles sharing halo rows (x and y neighbours, x fastest) run at the same
// time on the same die and meet their halo lines in that die's L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t w, uint32_t n) {
  const uint32_t q = n / kXcds, r = n % kXcds;  // XCDs below r take q + 1 tiles
  const uint32_t c = w % kXcds, slot = w / kXcds;
  return c < r ? c * (q + 1) + slot : r * (q + 1) + (c - r) * q + slot;
}

// Vector-memory counter of gfx9-family parts: it counts loads and stores, in
// issue order.  Step one of the next plane needs only the loads prefetched
// before this plane's step-two stores, i.e. a wait until at most Q operations
// remain.  The compiler takes the strictest wait over the paths that reach the
// loop head, and the paths that issue no stores (the march's entry, the waves
// without step-two work) would force vmcnt(0), i.e. waiting for the stores to
// reach memory as well.  Those paths wait for their prefetched loads here, with
// an empty asm that reads the destination registers: the compiler places the
// wait itself and knows that nothing is pending on that path afterwards.  (An
// explicit s_waitcnt is outside its model: it kept the loads pending after one,
// and the loop head's merged wait was vmcnt(0).)
template <class V, int N>
__device__ __forceinline__ void settle_loads(const V (&pre)[N], const int& mask) {
#pragma unroll
  for (int q = 0; q < N; ++q) asm volatile("" ::"v"(pre[q]));
  asm volatile("" ::"v"(mask));
}

// LDS writes complete, then the workgroup barrier.  Unlike __syncthreads() this
// does not wait for outstanding global loads, so the next plane's prefetch
// stays in flight across the barrier.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Rows of the two-step kernel's LDS ring.  Step two on plane z pulls population q
// from plane z - c_z[q], so the c_z = -1 populations of a plane are read for one
// plane of the march, c_z = 0 for two and c_z = +1 for three; one more slot per
// set lets the next plane be written while this one is still read.
template <class L>
struct ZRing {
  static constexpr int count(int v) {
    int n = 0;
    for (int q = 0; q < L::Q; ++q) n += L::cz[q] == v ? 1 : 0;
    return n;
  }
  // position of q among the populations with the same c_z
  static constexpr int index(int q) {
    int n = 0;
    for (int r = 0; r < q; ++r) n += L::cz[r] == L::cz[q] ? 1 : 0;
    return n;
  }
  static constexpr int NA = count(-1), NB = count(0), NC = count(1);
  static constexpr int OFF_B = 2 * NA, OFF_C = OFF_B + 3 * NB;
  static constexpr int ROWS = OFF_C + 4 * NC;  // 57 for D3Q19 (= 3 Q), 81 for D3Q27
};

// Two fused steps per launch (temporal blocking).  A workgroup owns a TX x TY
// column of output cells and marches z over `zc` planes.  Step-one states of a
// (TX+2) x (TY+2) tile (one cell per lane, read from HBM) are kept in LDS for
// three consecutive planes, in the storage format, so step two sees exactly the
// values a second launch would read; step two pulls from LDS, so a cell's
// populations cross HBM once per two steps plus the tile halo.  The step-one
// loads of the next plane are issued before the LDS phase of the current one.
//
// `tile` is the workgroup's column tile and z chunk (x tiles fastest).  BoxWalls:
// the solid cells are the five stationary walls of a cubic lid-driven cavity
// (x = 0, x = nx-1, y = 0, z = 0, z = nz-1), decided from the coordinates
// instead of read from a mask.
template <class L, int S, int COLL, int FEAT, int BCK, int TX, int TY, bool BoxWalls = false>
__device__ __forceinline__ void lbm2_body(const KArgs<typename Store<S>::C>& a, int zc, uint32_t tile) {
  using ST = Store<S>;
  using T = typename ST::T;
  using C = typename ST::C;
  constexpr int Q = L::Q;
  constexpr int EX = TX + 2, EY = TY + 2, EC = EX * EY;
  constexpr bool kAll = (FEAT & FEAT_BC) != 0;
  constexpr bool kForce = (FEAT & FEAT_FORCE) != 0;
  static_assert((FEAT & FEAT_MLP) == 0, "the closure runs in the one-step kernel");
  static_assert(!kForce || !BoxWalls, "no force sample on the cavity walls");
  // step-one states, one row of EC tile cells per (set, slot, population): the
  // c_z = -1 / 0 / +1 populations of a plane are pulled while it is the newest /
  // middle / oldest of three, and each set has one slot more than it has planes
  // in use, so step one of the next plane can be written while slower waves
  // still pull for this one: one barrier per plane
  using R = ZRing<L>;
  __shared__ T s[R::ROWS * EC];
  // solid flags of the same tile cells (three planes): step two takes them from
  // here rather than waiting on a mask load of its own each plane
  __shared__ uint8_t s_solid[kAll ? 3 * EC : 1];
  // f_eq(1, bc_u), built once per workgroup with the code the one-step kernel runs
  __shared__ C s_eq[kAll ? Q : 1];
  const int t = threadIdx.x;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t ntx = uint32_t((nx + TX - 1) / TX), nty = uint32_t((ny + TY - 1) / TY);
  const int x0 = int(tile % ntx) * TX, y0 = int((tile / ntx) % nty) * TY, z0 = int(tile / (ntx * nty)) * zc;
  const int zend = z0 + zc < nz ? z0 + zc : nz;
  const T* fin = static_cast<const T*>(a.fin);
  // step-one cell of this lane (tile-local lx, ly; global gx, gy)
  const bool h1 = t < EC;
  const int lx = h1 ? t % EX : 0, ly = h1 ? t / EX : 0;
  const int gx0 = x0 - 1 + lx, gy0 = y0 - 1 + ly;
  const int gx = (gx0 % nx + nx) % nx;  // halo lanes of a tile wider than the lattice wrap more than once
  const int gy = (gy0 % ny + ny) % ny;
  // inlet / outlet cells read their neighbouring column's pulled state in step
  // one: gather there instead (see BCN_IN_F)
  bool nb_edge = false;
  if constexpr (kAll) {
    const int bck = BCK == BCK_RUNTIME ? a.bc_kind : BCK;
    nb_edge = (bck == TLBM_BC_CHANNEL_X && (gx == 0 || gx == nx - 1)) ||
              (bck == TLBM_BC_FAR_FIELD && gx == nx - 1 && !(a.bc_faces & TLBM_FACE_XP_PERIODIC));
  }
  const int gxn = nb_edge ? (gx == 0 ? 1 : nx - 2) : gx;
  // output cell of this lane.  The output lanes fill whole waves; lanes of a
  // ragged tile whose cell lies past the grid take the tile's last in-grid cell
  // instead and compute it from the same inputs with the same code, so every
  // wave runs the update unmasked; only lanes that own their cell store (`own`),
  // so no cell is written twice.
  static_assert(TX * TY % 64 == 0, "output lanes fill whole waves");
  const bool out_wave = __builtin_amdgcn_readfirstlane(t) < TX * TY;
  const bool h2 = t < TX * TY;
  const int ux_ = min(x0 + (h2 ? t % TX : 0), nx - 1), uy_ = min(y0 + (h2 ? t / TX : 0), ny - 1);
  const int ox = ux_ - x0, oy = uy_ - y0;
  const bool own = h2 && x0 + t % TX < nx && y0 + t / TX < ny;
  double unused_force = 0.0;
  // Ladd force of both steps (obstacles.py:91-130), each cell of the launch's output range
  // sampled once per step: step one by the lanes whose halo-tile cell is an interior cell of
  // this column tile inside the grid (no wrapped copy) on the chunk's own planes, step two by
  // the lanes that own their output cell.  One FP64 atomic per wave after the march.
  double fx = 0.0, fy = 0.0, fz = 0.0;
  bool sampled = false;
  const bool own1 = h1 && lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && gx0 < nx && gy0 < ny;

  if constexpr (kAll) {
    if (t == 0) {
      C e[Q];
      equilibrium_here<L>(C(1), a.bc_u, e);
#pragma unroll
      for (int q = 0; q < Q; ++q) s_eq[q] = e[q];
    }
    __syncthreads();
  }
  auto eq_unit = [&](C (&e)[Q]) {
#pragma unroll
    for (int q = 0; q < Q; ++q) e[q] = s_eq[q];
  };

  C pre[Q];  // pulled populations of this lane's step-one cell on the next plane
  int pre_mask = 0;  // its solid byte, compared only where step one uses it
  auto prefetch = [&](int p) {
    if (!h1) return;
    const int pz = p < 0 ? p + nz : (p >= nz ? p - nz : p);
    // the mask byte first: the loads complete in order, so consuming it does not
    // wait for the population loads behind it
    if constexpr (kAll && BoxWalls) {
      pre_mask = gx == 0 || gx == nx - 1 || gy == 0 || pz == 0 || pz == nz - 1;
    } else if constexpr (kAll) {
      if (a.solid != nullptr) pre_mask = a.solid[(uint32_t(pz) * uint32_t(ny) + uint32_t(gy)) * uint32_t(nx) + uint32_t(gx)];
    }
    gather<L, S>(pre, fin, a, gxn, gy, pz);
  };
  // first row of the set holding population q, for the plane written j-th in the march
  auto set_row = [](int cz, int j) {
    return cz < 0 ? (j & 1) * R::NA : cz == 0 ? R::OFF_B + (j % 3) * R::NB : R::OFF_C + (j & 3) * R::NC;
  };
  auto step_one = [&](int p, int j) {
    if (!h1) return;
    const int pz = p < 0 ? p + nz : (p >= nz ? p - nz : p);
    const uint32_t cell = (uint32_t(pz) * uint32_t(ny) + uint32_t(gy)) * uint32_t(nx) + uint32_t(gx);
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = pre[q];
    auto unused = [](int, int, int, int) { return C(0); };
    const bool solid = pre_mask != 0;
    if constexpr (kForce) {
      double sx = 0.0, sy = 0.0, sz = 0.0;
      update_cell<L, S, COLL, true, FEAT, BCN_IN_F, BCK>(f, a, gx, gy, pz, cell, solid, nb_edge, sx, sy, sz, unused,
                                                         eq_unit);
      if (own1 && solid && p >= z0 && p < zend) {
        fx += sx;
        fy += sy;
        fz += sz;
        sampled = true;
      }
    } else {
      update_cell<L, S, COLL, true, FEAT, BCN_IN_F, BCK>(f, a, gx, gy, pz, cell, solid, nb_edge, unused_force,
                                                         unused_force, unused_force, unused, eq_unit);
    }
    const int row_a = set_row(-1, j), row_b = set_row(0, j), row_c = set_row(1, j);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int row = (L::cz[q] < 0 ? row_a : L::cz[q] == 0 ? row_b : row_c) + R::index(q);
      s[row * EC + t] = put_out<L, S, COLL>(f[q], q);
    }
    if constexpr (kAll) s_solid[(j % 3) * EC + t] = solid ? 1 : 0;
  };

  // The march: iteration k computes step one on plane z0 + k + 1 (the (k + 2)-th
  // plane written) and, from k = 0 on, step two on plane z = z0 + k.  Between two
  // barriers a wave may run step two of iteration k while another runs step one
  // of k + 1: the rows that step one writes (set slots (k + 3) mod 2 / 3 / 4) are
  // not among those step two of k reads (k mod 2, (k + 1) mod 3, k mod 4), and no
  // wave passes the next barrier before all are done.
  //
  // Waits for the prefetched loads: every path into the main loop's head either
  // issues its step-two stores after the loads or waits for the loads itself
  // (the fill planes, the waves without output cells), so that the head need
  // not wait for the previous plane's stores.  The compiler still merges the
  // head's wait to vmcnt(0) in the channel kernel; unmasked stores, step one in
  // the prefetched registers and scalar-base stores relaxed it to the loads
  // alone but left the headline's time unchanged (68.9 GLUPS either way), so the
  // simpler form stays.
  prefetch(z0 - 1);
  for (int k = -2; k < 0; ++k) {  // the two fill planes: step one only
    const int p1 = z0 + k + 1;
    step_one(p1, k + 2);
    prefetch(p1 + 1);  // p1 + 1 <= z0 + 1 <= zend
    lds_barrier();
  }
  settle_loads(pre, pre_mask);
  for (int k = 0; z0 + k < zend; ++k) {
    const int z = z0 + k;
    const int p1 = z + 1;
    step_one(p1, k + 2);
    if (p1 < zend) prefetch(p1 + 1);  // the next iteration's step-one plane
    lds_barrier();
    if (out_wave) {
      // step-one value of population q pulled into cell (x, y) of plane z (same
      // plane only: the boundary overrides read along x), written as plane
      // k + 1 - c_z[q] of the march
      const int row_a = set_row(-1, k + 2), row_b = set_row(0, k + 1), row_c = set_row(1, k);
      auto lds_at = [&](int q, int x, int y, int) -> C {
        const int row = (L::cz[q] < 0 ? row_a : L::cz[q] == 0 ? row_b : row_c) + R::index(q);
        const int sx = x - L::cx[q] - (x0 - 1), sy = y - L::cy[q] - (y0 - 1);
        return ST::template get<L>(s[row * EC + sy * EX + sx], q);
      };
      C f[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) f[q] = lds_at(q, ux_, uy_, z);
      const uint32_t cell = (uint32_t(z) * uint32_t(ny) + uint32_t(uy_)) * uint32_t(nx) + uint32_t(ux_);
      bool solid = false;
      if constexpr (kAll) solid = s_solid[((k + 1) % 3) * EC + (oy + 1) * EX + (ox + 1)] != 0;
      if constexpr (kForce) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        update_cell<L, S, COLL, true, FEAT, BCN_PULL, BCK>(f, a, ux_, uy_, z, cell, solid, false, sx, sy, sz, lds_at,
                                                           eq_unit);
        if (own && solid) {
          fx += sx;
          fy += sy;
          fz += sz;
          sampled = true;
        }
      } else {
        update_cell<L, S, COLL, true, FEAT, BCN_PULL, BCK>(f, a, ux_, uy_, z, cell, solid, false, unused_force,
                                                           unused_force, unused_force, lds_at, eq_unit);
      }
      T* fout = static_cast<T*>(a.fout);
      const uint32_t off = uint32_t(uy_) * uint32_t(a.osy) + uint32_t(ux_);
      // cavity rows (pitch = edge) rarely start on a cache line, and non-temporal
      // stores of the partial lines at a tile's ends each reach HBM: on 64
      // cavities of edges 32-160 that wrote 1.21x the least bytes and ran at 42.1
      // GLUPS; through L2, 1.02x and 48.6 GLUPS
      if (own) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          T* const base = fout + int64_t(q) * a.osq + int64_t(z) * a.osz;
          if constexpr (BoxWalls) st_l2(base, off, put_out<L, S, COLL>(f[q], q));
          else st_nt(base, off, put_out<L, S, COLL>(f[q], q));
        }
      }
    } else {
      settle_loads(pre, pre_mask);
    }
  }
  if constexpr (kForce) {
    if (a.force != nullptr) force_flush(a.force, fx, fy, fz, sampled);
  }
}

template <class L, int S, int COLL, bool STREAM, int FEAT, int BCK>
__global__ __launch_bounds__(256) void k_lbm(const KArgs<typename Store<S>::C> a) {
  lbm_body<L, S, COLL, STREAM, FEAT, BCK>(a);
}

// two-step tile: TX x TY output cells, (TX+2)(TY+2) lanes rounded to whole waves
template <int TX, int TY>
constexpr int tb_threads() {
  return ((TX + 2) * (TY + 2) + 63) / 64 * 64;
}

template <class L, int S, int COLL, int FEAT, int BCK, int TX, int TY>
__global__ __launch_bounds__((tb_threads<TX, TY>())) void k_lbm2(const KArgs<typename Store<S>::C> a, int zc) {
  lbm2_body<L, S, COLL, FEAT, BCK, TX, TY>(a, zc, xcd_tile(blockIdx.x, gridDim.x));
}

// The learned-closure step.  Its collision is long (the network runs on the
// matrix cores beside ~280 VALU instructions per wave), and at the 4 waves per
// SIMD its registers allow, a wave that loads, collides and stores one cell
// leaves the SIMD idle while its own loads are in flight.  So a workgroup
// marches `zper` planes of one tile: the next plane's pulled populations (and
// solid byte) are loaded while the current plane is collided, and the network's
// operand blob is staged once into LDS, whose reads count on the LDS counter
// and so never wait behind the prefetch as global reads of the blob would.
template <class L, int S, int COLL, bool STREAM, int FEAT, int BCK>
__device__ __forceinline__ void lbm_mlp_body(const KArgs<typename Store<S>::C>& a, int zper) {
  static_assert(STREAM, "the learned closure runs in the fused step");
  using ST = Store<S>;
  using T = typename ST::T;
  using C = typename ST::C;
  constexpr int Q = L::Q;
  constexpr bool kAll = (FEAT & FEAT_BC) != 0;
  // with 16-bit deviation storage the step runs in f - w from load to store: the
  // loads skip adding w_q and the collision folds it into one term per cell
  constexpr bool kDevIn = dev_out<S, COLL>() && (FEAT & FEAT_MLP) != 0;
  __shared__ mlp::Packed s_blob;
  // f_eq(rho = 1, bc_u) for the inlet and the far-field faces, formed once per
  // workgroup instead of by every wave with such a lane on every plane, and its
  // deviation f_eq - w for the step in deviations
  __shared__ C s_eq[kAll ? Q : 1];
  __shared__ C s_eqd[kAll && kDevIn ? Q : 1];
  {
    const uint32_t* src = static_cast<const uint32_t*>(a.mlp);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&s_blob);
    const int t = int(threadIdx.y * blockDim.x + threadIdx.x);
    for (int i = t; i < int(sizeof(mlp::Packed) / 4); i += int(blockDim.x * blockDim.y)) dst[i] = src[i];
    if constexpr (kAll) {
      if (t == 0) {
        C e[Q];
        equilibrium_here<L>(C(1), a.bc_u, e);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          s_eq[q] = e[q];
          if constexpr (kDevIn) s_eqd[q] = e[q] - C(L::w[q]);
        }
      }
    }
  }
  __syncthreads();

  const uint32_t tile = blockIdx.x;
  const uint32_t txi = tile % uint32_t(a.tiles_x);
  const uint32_t rest = tile / uint32_t(a.tiles_x);
  const uint32_t tyi = rest % uint32_t(a.tiles_y);
  const int z0 = int(rest / uint32_t(a.tiles_y)) * zper;
  const int z1 = min(z0 + zper, a.nz);
  const int x = int(txi * blockDim.x + threadIdx.x);
  const int y = int(tyi * blockDim.y + threadIdx.y);
  const bool active = x < a.nx && y < a.ny;
  // Waves wholly past the grid's edge (tiles wider or taller than the grid) leave;
  // there is no barrier after this point.
  if (__all(!active)) return;
  // Lanes of a ragged tile past the grid's edge work on the last in-grid cell of
  // their row and store the same bits as the lane that owns it (which is in the
  // same wave), so every store is unmasked: the compiler merges its wait counts
  // over the paths into the loop head, and a path with masked or skipped stores
  // makes that head wait for the previous plane's stores as well as for the
  // prefetch.  Only the force sample, a sum, is taken from the owning lanes alone.
  const int xx = min(x, a.nx - 1);
  const int yy = min(y, a.ny - 1);
  const T* fin = static_cast<const T*>(a.fin);
  const int8_t* solid_mask = kAll ? a.solid : nullptr;

  T pre[Q];
  int pre_mask = 0;
  auto prefetch = [&](int z, int px, int py) {
    gather_raw<L, S>(pre, fin, a, px, py, z);
    if (solid_mask != nullptr)
      pre_mask = solid_mask[(uint32_t(z) * uint32_t(a.ny) + uint32_t(py)) * uint32_t(a.nx) + uint32_t(px)];
  };
  prefetch(z0, xx, yy);
  // the lane's store offset within a plane, fixed for the march (computed in the
  // loop, its 32-bit multiply-add became a 64-bit one whose unused high operand
  // register could be a prefetch destination, i.e. a wait for the prefetch)
  uint32_t off = uint32_t(yy) * uint32_t(a.osy) + uint32_t(xx);
  asm volatile("" : "+v"(off));
  // the loop head then waits only for the prefetch, with the previous plane's
  // stores still in flight
  settle_loads(pre, pre_mask);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  bool any_solid = false;
  // The inlet (x = 0) and outlet (x = nx - 1) lanes need the pulled state of
  // x = 1 and x = nx - 2, which a lane of the same wave has just gathered.  Before
  // the update such a lane takes that state into its own f with a lane permute,
  // and `pulled` hands back f itself (the inlet sums it before overwriting f, the
  // outlet copies it onto itself), instead of Q more loads and their address
  // arithmetic.  The inlet's neighbour is always lane 1 of its wave (tiles and
  // waves start at multiples of 64 in x); the outlet's is in the same wave unless
  // nx - 1 is a multiple of 64, and is then loaded.  Clamped lanes past the edge
  // act as outlet lanes too, hence a fixed source lane, not lane - 1.  With a
  // force buffer the loads stay: the force sample reads a solid edge cell's own
  // populations before the boundary.
  constexpr bool kForceFeat = (FEAT & FEAT_FORCE) != 0;
  const bool out_shfl = !kForceFeat && ((a.nx - 1) & 63) != 0;
  struct EqTables {
    const C* e;
    const C* ed;
    __device__ void operator()(C (&o)[Q]) const {
#pragma unroll
      for (int q = 0; q < Q; ++q) o[q] = e[q];
    }
    __device__ void dev(C (&o)[Q]) const {
#pragma unroll
      for (int q = 0; q < Q; ++q) o[q] = ed[q];
    }
  };
  const EqTables eq_unit{s_eq, s_eqd};
  for (int z = z0; z < z1; ++z) {
    // Nothing per lane is carried across planes but the prefetched values: the
    // blob's LDS reads stay inside the loop (the memory clobber), and the lane's
    // coordinates are re-derived from an opaque copy, so the Q gather offsets
    // and the boundary's pull offsets are recomputed per plane (a few VALU
    // instructions) instead of held live across the march, where they spilled.
    asm volatile("" ::: "memory");
    int xl = xx, yl = yy;
    asm volatile("" : "+v"(xl), "+v"(yl));
    C f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if constexpr (kDevIn) f[q] = ST::get_dev(pre[q]);
      else f[q] = ST::template get<L>(pre[q], q);
    }
    const bool solid = pre_mask != 0;
    any_solid |= solid;
    if constexpr (kAll && !kForceFeat) {
      const int bck = BCK == BCK_RUNTIME ? a.bc_kind : BCK;
      const bool take = take_neighbour(bck, a.bc_faces, xl, a.nx, out_shfl);
      if (__any(take)) {
        const int src = xl == 0 ? 1 : ((a.nx - 2) & 63);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const C g = __shfl(f[q], src, 64);
          f[q] = take ? g : f[q];
        }
      }
    }
    auto pulled = [&](int q, int px, int py, int pz) -> C {
      if (!kForceFeat && (px == 1 || out_shfl)) return f[q];
      const C v = pull_one<ST, L, C, true>(fin, a, q, px, py, pz);
      return kDevIn ? v - C(L::w[q]) : v;
    };
    const uint32_t cell = (uint32_t(z) * uint32_t(a.ny) + uint32_t(yl)) * uint32_t(a.nx) + uint32_t(xl);
    // the next plane's loads go out after the boundary's own pulls (inlet /
    // outlet lanes), so waiting for those never waits for the prefetch
    auto next = [&] {
      if (z + 1 < z1) prefetch(z + 1, xl, yl);
    };
    update_cell<L, S, COLL, STREAM, FEAT, BCN_PULL, BCK, kDevIn>(f, a, xl, yl, z, cell, solid, false, fx, fy, fz,
                                                                  pulled, eq_unit, (mlp::LdsPacked)(&s_blob), next);
    T* fout = static_cast<T*>(a.fout);
    // re-derived per plane, so that the stores keep the scalar-base + 32-bit
    // offset form (a 64-bit offset hoisted out of the loop costs one 64-bit add
    // per store)
    uint32_t offl = off;
    asm volatile("" : "+v"(offl));
#pragma unroll
    for (int q = 0; q < Q; ++q)
      st_nt(fout + int64_t(q) * a.osq + int64_t(z) * a.osz, offl, put_out<L, S, COLL>(f[q], q));
  }
  if constexpr ((FEAT & FEAT_FORCE) != 0) {
    if (a.force != nullptr)
      force_flush(a.force, active ? fx : 0.0, active ? fy : 0.0, active ? fz : 0.0, active && any_solid);
  }
}

// planes marched per workgroup by the learned-closure step
#ifndef TLBM_MLP_ZPER
#define TLBM_MLP_ZPER 16
#endif

// The march pays where the step is bound by instruction issue, i.e. with 16-bit
// storage (256^3 channel + sphere, fp16 deviations, one box: 43.9 GLUPS one cell
// per lane, 47.0 marching 8 planes, 47.1 marching 16; on another, 2 and 4 planes
// were slower than 8).  With fp32 storage the step moves twice the bytes and one
// cell per lane was as fast or faster (36.7 against 35.4 / 34.9 / 33.5 for 4 / 8 /
// 16 planes); fp64 would need 2Q more registers than 4 waves per SIMD allow.
template <int S>
constexpr bool mlp_marches() {
  return S == ST_F16 || S == ST_BF16 || S == ST_F16_DEV || S == ST_BF16_DEV;
}

template <class L, int S, int COLL, bool STREAM, int FEAT, int BCK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TLBM_MLP_WAVES))) void k_lbm_mlp(
    const KArgs<typename Store<S>::C> a, int zper) {
  if constexpr (mlp_marches<S>()) {
    lbm_mlp_body<L, S, COLL, STREAM, FEAT, BCK>(a, zper);
  } else {
    (void)zper;
    lbm_body<L, S, COLL, STREAM, FEAT, BCK>(a);
  }
}

// ---------------------------------------------------------------------------
// elementwise ops (contiguous fields)
// ---------------------------------------------------------------------------

template <class L, typename T>
__global__ __launch_bounds__(256) void k_equilibrium(const T* rho, const T* ux, const T* uy, const T* uz,
                                                     T* out, uint64_t N) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T feq[L::Q];
  equilibrium<L>(rho[i], ux[i], uy[i], uz ? uz[i] : T(0), feq);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) out[uint64_t(q) * N + i] = feq[q];
}

template <class L, typename T>
__global__ __launch_bounds__(256) void k_macroscopic(const T* f, T* rho, T* ux, T* uy, T* uz, uint64_t N) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T v[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) v[q] = f[uint64_t(q) * N + i];
  const Macro<T> m = macro<L>(v);
  rho[i] = m.rho;
  ux[i] = m.ux;
  uy[i] = m.uy;
  if (uz) uz[i] = m.uz;
}

template <class L, typename T>
__global__ __launch_bounds__(256) void k_bounce_back(const T* fin, T* fout, const int8_t* mask, uint64_t N) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T v[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) v[q] = fin[uint64_t(q) * N + i];
  const bool s = mask[i] != 0;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) fout[uint64_t(q) * N + i] = s ? v[L::opp[q]] : v[q];
}

// ---------------------------------------------------------------------------
// boundary conditions as separate ops (eager forms of the fused boundaries)
// ---------------------------------------------------------------------------

// One cell of a post-stream D3Q19 state, out of place: the face conditions of `kind`, then full-way
// bounce-back where mask_a is set, then where mask_b is (a cell in both swaps twice).  Every value
// read comes from the input, i.e. the state before any condition: the inlet's density from x = 1
// and the outlet's copy of x = nx - 2 are taken before the lateral far-field faces are written, as
// in boundaries3d.py:407-443 / 446-537 (nx >= 3 is checked on the host).
template <typename T>
__global__ __launch_bounds__(256) void k_boundary(const T* fin, T* fout, int nz, int ny, int nx, int kind, int faces,
                                                  T bux, T buy, T buz, const int8_t* mask_a, const int8_t* mask_b) {
  using L = D3Q19;
  constexpr int Q = L::Q;
  const uint64_t N = uint64_t(nz) * ny * nx;
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int x = int(i % uint64_t(nx));
  const int y = int((i / uint64_t(nx)) % uint64_t(ny));
  const int z = int(i / (uint64_t(nx) * ny));
  T f[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = fin[uint64_t(q) * N + i];
  if (kind == TLBM_BC_CHANNEL_X) {
    if (x == 0) {  // equilibrium inlet, u = (u_in, 0, 0), density of x = 1
      T rho1 = T(0);
#pragma unroll
      for (int q = 0; q < Q; ++q) rho1 += fin[uint64_t(q) * N + i + 1];
      equilibrium<L>(rho1, bux, T(0), T(0), f);
    } else if (x == nx - 1) {  // zero-gradient outlet
#pragma unroll
      for (int q = 0; q < Q; ++q) f[q] = fin[uint64_t(q) * N + i - 1];
    }
  } else if (kind == TLBM_BC_FAR_FIELD) {
    const bool lateral = (y == 0 && (faces & TLBM_FACE_YM)) || (y == ny - 1 && (faces & TLBM_FACE_YP)) ||
                         (z == 0 && (faces & TLBM_FACE_ZM)) || (z == nz - 1 && (faces & TLBM_FACE_ZP));
    if (lateral || (x == 0 && !(faces & TLBM_FACE_XM_PERIODIC))) {
      equilibrium<L>(T(1), bux, buy, buz, f);  // free stream, rho = 1
    } else if (x == nx - 1 && !(faces & TLBM_FACE_XP_PERIODIC)) {
#pragma unroll
      for (int q = 0; q < Q; ++q) f[q] = fin[uint64_t(q) * N + i - 1];
    }
  } else if (kind == TLBM_BC_LID_Y) {
    if (y == ny - 1) zou_he_lid(f, bux);  // every cell of the lid plane, corners included
  }
  const bool swap = (mask_a != nullptr && mask_a[i] != 0) != (mask_b != nullptr && mask_b[i] != 0);
#pragma unroll
  for (int q = 0; q < Q; ++q) fout[uint64_t(q) * N + i] = swap ? f[L::opp[q]] : f[q];
}

// sphere_mask (boundaries3d.py:29-49) in float32 as the reference evaluates it: centre and
// coordinates rounded to float32, each square and each sum rounded in turn, against radius^2
// formed in double and rounded to float32
__global__ __launch_bounds__(256) void k_sphere_mask(int nx, int ny, int nz, float cx, float cy, float cz, float r2,
                                                     uint8_t* out) {
#pragma clang fp contract(off)
  const uint64_t n = uint64_t(nz) * ny * nx;
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float dx = float(int(i % uint64_t(nx))) - cx;
  const float dy = float(int((i / uint64_t(nx)) % uint64_t(ny))) - cy;
  const float dz = float(int(i / (uint64_t(nx) * ny))) - cz;
  const float d = dx * dx + dy * dy + dz * dz;
  out[i] = d <= r2 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// gradient eddy-viscosity closures (WALE, Vreman) -> per-cell tau_eff
// ---------------------------------------------------------------------------

// velocity component of cell i, or the wall velocity on solid cells
template <typename T>
__device__ __forceinline__ T vel_at(const T* u, const int8_t* solid, T wall, uint64_t i) {
  return (solid != nullptr && solid[i] != 0) ? wall : u[i];
}

// d/dx along one axis of length n at position p (element stride st): centred inside,
// second-order one-sided on the end planes, forward difference on an axis of length 2,
// zero on an axis of length 1 (turbulence.py:595-640)
template <typename T>
__device__ __forceinline__ T edge_derivative(const T* u, const int8_t* solid, T wall, uint64_t i, int p, int n,
                                             uint64_t st) {
  if (n == 1) return T(0);
  if (n == 2) {
    const uint64_t i0 = i - uint64_t(p) * st;
    return vel_at(u, solid, wall, i0 + st) - vel_at(u, solid, wall, i0);
  }
  if (p == 0)
    return T(0.5) * (T(-3) * vel_at(u, solid, wall, i) + T(4) * vel_at(u, solid, wall, i + st) -
                     vel_at(u, solid, wall, i + 2 * st));
  if (p == n - 1)
    return T(0.5) * (T(3) * vel_at(u, solid, wall, i) - T(4) * vel_at(u, solid, wall, i - st) +
                     vel_at(u, solid, wall, i - 2 * st));
  return T(0.5) * (vel_at(u, solid, wall, i + st) - vel_at(u, solid, wall, i - st));
}

// tau_eff = clip(tau + 3 nu_t, 0.5001, max(1, tau + 0.5)) (turbulence.py:875-888) with nu_t of
// WALE (turbulence.py:754-841) or Vreman (turbulence.py:844-872), lattice filter width 1
template <typename T>
__global__ __launch_bounds__(256) void k_gradient_tau(const T* ux, const T* uy, const T* uz, const int8_t* solid,
                                                      T wx, T wy, T wz, int nz, int ny, int nx, int model, T coef,
                                                      T tau, T tau_cap, T* tau_out) {
  const uint64_t n = uint64_t(nz) * ny * nx;
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = int(i % uint64_t(nx));
  const int y = int((i / uint64_t(nx)) % uint64_t(ny));
  const int z = int(i / (uint64_t(nx) * ny));
  const T* u[3] = {ux, uy, uz};
  const T w[3] = {wx, wy, wz};
  const int pos[3] = {x, y, z};
  const int len[3] = {nx, ny, nz};
  const uint64_t st[3] = {1, uint64_t(nx), uint64_t(nx) * ny};
  T g[3][3];  // g[i][j] = du_i / dx_j
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) g[c][d] = edge_derivative(u[c], solid, w[c], i, pos[d], len[d], st[d]);
  T nu;
  if (model == TLBM_SGS_WALE) {
    T g2[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) g2[r][c] = g[r][0] * g[0][c] + g[r][1] * g[1][c] + g[r][2] * g[2][c];
    const T tr = (g2[0][0] + g2[1][1] + g2[2][2]) / T(3);
    const T d01 = T(0.5) * (g2[0][1] + g2[1][0]), d02 = T(0.5) * (g2[0][2] + g2[2][0]),
            d12 = T(0.5) * (g2[1][2] + g2[2][1]);
    T sd = (g2[0][0] - tr) * (g2[0][0] - tr) + (g2[1][1] - tr) * (g2[1][1] - tr) + (g2[2][2] - tr) * (g2[2][2] - tr) +
           T(2) * (d01 * d01 + d02 * d02 + d12 * d12);
    const T s01 = T(0.5) * (g[0][1] + g[1][0]), s02 = T(0.5) * (g[0][2] + g[2][0]), s12 = T(0.5) * (g[1][2] + g[2][1]);
    T sn = g[0][0] * g[0][0] + g[1][1] * g[1][1] + g[2][2] * g[2][2] + T(2) * (s01 * s01 + s02 * s02 + s12 * s12);
    sd = sd > T(0) ? sd : T(0);
    sn = sn > T(0) ? sn : T(0);
    const T rsd = sqrt(sd), rsn = sqrt(sn);
    nu = coef * coef * (sd * rsd) / (sn * sn * rsn + sd * sqrt(rsd) + T(1e-30));
  } else {
    T b[3][3];  // beta = alpha^T alpha
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) b[r][c] = g[0][r] * g[0][c] + g[1][r] * g[1][c] + g[2][r] * g[2][c];
    T a2 = T(0);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) a2 += g[r][c] * g[r][c];
    T bb = b[0][0] * b[1][1] - b[0][1] * b[0][1] + b[0][0] * b[2][2] - b[0][2] * b[0][2] + b[1][1] * b[2][2] -
           b[1][2] * b[1][2];
    bb = bb > T(0) ? bb : T(0);
    nu = coef * sqrt(bb / (a2 + T(1e-30)));
  }
  T te = tau + T(3) * nu;
  te = te < T(0.5001) ? T(0.5001) : (te > tau_cap ? tau_cap : te);
  tau_out[i] = te;
}

// ---------------------------------------------------------------------------
// layout-aware FP64 reductions and scaling
// ---------------------------------------------------------------------------

struct Layout {
  int64_t sq, sz, sy;
  int nx, ny, nz;
};

// sums of f, c_x f, c_y f, c_z f over all cells (solid != nullptr: only solid
// cells, Ladd factor 2, out[0..2] = force)
template <class L, int S>
__global__ __launch_bounds__(256) void k_moment_sums(const void* fv, Layout lay, const int8_t* solid, double* out) {
  using ST = Store<S>;
  const typename ST::T* f = static_cast<const typename ST::T*>(fv);
  double s0 = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  const uint64_t N = uint64_t(lay.nx) * uint64_t(lay.ny) * uint64_t(lay.nz);
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += stride) {
    if (solid != nullptr && solid[i] == 0) continue;
    const uint64_t x = i % uint64_t(lay.nx);
    const uint64_t r = i / uint64_t(lay.nx);
    const uint64_t y = r % uint64_t(lay.ny);
    const uint64_t z = r / uint64_t(lay.ny);
    const int64_t base = int64_t(z) * lay.sz + int64_t(y) * lay.sy + int64_t(x);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      const double v = double(ST::template get<L>(f[base + int64_t(q) * lay.sq], q));
      s0 += v;
      if (L::cx[q] != 0) sx += double(L::cx[q]) * v;
      if (L::cy[q] != 0) sy += double(L::cy[q]) * v;
      if (L::cz[q] != 0) sz += double(L::cz[q]) * v;
    }
  }
  s0 = wave_sum(s0);
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  __shared__ double part[4][4];
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  if (lane == 0) {
    part[wid][0] = s0;
    part[wid][1] = sx;
    part[wid][2] = sy;
    part[wid][3] = sz;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < int(blockDim.x / 64); ++w) t += part[w][threadIdx.x];
    if (solid == nullptr) {
      atomicAdd(out + threadIdx.x, t);
    } else if (threadIdx.x > 0) {
      atomicAdd(out + threadIdx.x - 1, 2.0 * t);
    }
  }
}

// sum over all cells of the clamped Smagorinsky tau_eff (turbulence.py:183-208) of
// the local non-equilibrium stress, f_eq built with rho (the eager ops' form)
template <class L, int S>
__global__ __launch_bounds__(256) void k_tau_sum(const void* fv, Layout lay, double tau, double cs_delta_sq,
                                                 double* out) {
  using ST = Store<S>;
  using C = typename ST::C;
  const typename ST::T* f = static_cast<const typename ST::T*>(fv);
  double acc = 0.0;
  const uint64_t N = uint64_t(lay.nx) * uint64_t(lay.ny) * uint64_t(lay.nz);
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  const C tau_cap = C(tau + 0.5 > 1.0 ? tau + 0.5 : 1.0);
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += stride) {
    const uint64_t x = i % uint64_t(lay.nx);
    const uint64_t r = i / uint64_t(lay.nx);
    const uint64_t y = r % uint64_t(lay.ny);
    const uint64_t z = r / uint64_t(lay.ny);
    const int64_t base = int64_t(z) * lay.sz + int64_t(y) * lay.sy + int64_t(x);
    C v[L::Q], feq[L::Q], pi[6];
#pragma unroll
    for (int q = 0; q < L::Q; ++q) v[q] = ST::template get<L>(f[base + int64_t(q) * lay.sq], q);
    const Macro<C> mc = macro<L>(v);
    equilibrium<L>(mc.rho, mc.ux, mc.uy, mc.uz, feq);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) v[q] -= feq[q];
    neq_stress<L>(v, pi);
    acc += double(smagorinsky_tau(pi_norm(pi), mc.rho_safe, C(tau), C(cs_delta_sq), tau_cap));
  }
  acc = wave_sum(acc);
  __shared__ double part[4];
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  if (lane == 0) part[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < int(blockDim.x / 64); ++w) t += part[w];
    atomicAdd(out, t);
  }
}

template <class L, int S>
__global__ __launch_bounds__(256) void k_scale(void* fv, Layout lay, const double* scale) {
  using ST = Store<S>;
  typename ST::T* f = static_cast<typename ST::T*>(fv);
  const uint64_t N = uint64_t(lay.nx) * uint64_t(lay.ny) * uint64_t(lay.nz);
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint64_t x = i % uint64_t(lay.nx);
  const uint64_t r = i / uint64_t(lay.nx);
  const uint64_t y = r % uint64_t(lay.ny);
  const uint64_t z = r / uint64_t(lay.ny);
  const int64_t base = int64_t(z) * lay.sz + int64_t(y) * lay.sy + int64_t(x);
  const typename ST::C s = typename ST::C(*scale);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int64_t k = base + int64_t(q) * lay.sq;
    f[k] = ST::template put<L>(ST::template get<L>(f[k], q) * s, q);
  }
}

}  // namespace tlbm

// ===========================================================================
// C ABI
// ===========================================================================

using namespace tlbm;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace

namespace tlbm {
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace tlbm

namespace {

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(TLBM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return TLBM_OK;
}

bool lattice_ok(int l) { return l == TLBM_D2Q9 || l == TLBM_D3Q19 || l == TLBM_D3Q27; }
bool store_ok(int s) { return s >= TLBM_STORE_F32 && s <= TLBM_STORE_BF16_DEV; }

size_t store_size(int s) {
  switch (s) {
    case TLBM_STORE_F32:
      return 4;
    case TLBM_STORE_F64:
      return 8;
    default:
      return 2;
  }
}

int check_dims(int lattice, int nz, int ny, int nx) {
  if (!lattice_ok(lattice)) return fail(TLBM_ERR_INVALID, "unknown lattice Q=%d", lattice);
  if (nz < 1 || ny < 1 || nx < 1) return fail(TLBM_ERR_SHAPE, "grid dims must be >= 1 (got %d,%d,%d)", nz, ny, nx);
  if (lattice == TLBM_D2Q9 && nz != 1) return fail(TLBM_ERR_SHAPE, "D2Q9 requires nz == 1 (got %d)", nz);
  const uint64_t n = uint64_t(nz) * uint64_t(ny) * uint64_t(nx);
  if (n >= (uint64_t(1) << 31)) return fail(TLBM_ERR_SHAPE, "grid of %llu cells exceeds 2^31", (unsigned long long)n);
  return TLBM_OK;
}

// strides (q, z, y) in elements, x stride 1; NULL = contiguous [q][z][y][x]
int resolve_strides(const int64_t* s, int lattice, int nz, int ny, int nx, size_t esize, int64_t out[3],
                    const char* which) {
  if (s == nullptr) {
    out[0] = int64_t(nz) * ny * nx;
    out[1] = int64_t(ny) * nx;
    out[2] = nx;
  } else {
    out[0] = s[0];
    out[1] = s[1];
    out[2] = s[2];
  }
  if (out[0] <= 0 || out[2] < nx || (nz > 1 && out[1] <= 0))
    return fail(TLBM_ERR_SHAPE, "%s strides (%lld, %lld, %lld) are not a valid population layout", which,
                (long long)out[0], (long long)out[1], (long long)out[2]);
  // per-lane offsets y*sy + x must stay below 2^32 bytes
  const uint64_t span = (uint64_t(ny - 1) * uint64_t(out[2]) + uint64_t(nx)) * esize;
  if (span >= (uint64_t(1) << 32))
    return fail(TLBM_ERR_SHAPE, "%s: one z-plane spans %llu bytes; the per-lane offset limit is 4 GiB", which,
                (unsigned long long)span);
  (void)lattice;
  return TLBM_OK;
}

template <typename C>
int fill_args(KArgs<C>& a, int lattice, const void* fin, const int64_t* is, void* fout, const int64_t* os, int nz,
              int ny, int nx, const tlbm_collision* coll, const int8_t* solid, const tlbm_boundary* bc,
              double* force, int* bx_out, int* by_out, uint32_t* nblocks) {
  memset(&a, 0, sizeof(a));
  a.fin = fin;
  a.fout = fout;
  a.solid = solid;
  a.force = force;
  a.isq = is[0];
  a.isz = is[1];
  a.isy = is[2];
  a.osq = os[0];
  a.osz = os[1];
  a.osy = os[2];
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  const int bx = nx <= 64 ? 64 : (nx <= 128 ? 128 : 256);
  const int by = 256 / bx;
  a.tiles_x = (nx + bx - 1) / bx;
  a.tiles_y = (ny + by - 1) / by;
  *bx_out = bx;
  *by_out = by;
  const uint64_t nb = uint64_t(a.tiles_x) * uint64_t(a.tiles_y) * uint64_t(nz);
  if (nb >= (uint64_t(1) << 31)) return fail(TLBM_ERR_SHAPE, "too many workgroups");
  *nblocks = uint32_t(nb);
  a.tau_mode = TLBM_TAU_CONST;
  a.eq_floor = 1;
  a.omega = C(1);
  if (coll != nullptr) {
    if (coll->kind != TLBM_COLL_NONE && !(coll->tau > 0.5))
      return fail(TLBM_ERR_INVALID, "tau must be > 0.5 (got %g)", coll->tau);
    a.tau_mode = coll->tau_mode;
    a.eq_floor = coll->eq_rho_floor;
    a.tau = C(coll->tau);
    a.omega = C(1.0 / coll->tau);
    a.cs_delta_sq = C(coll->cs_delta_sq);
    a.tau_cap = C(coll->tau + 0.5 > 1.0 ? coll->tau + 0.5 : 1.0);
    for (int i = 0; i < 8; ++i) a.rate[i] = C(coll->rate[i]);
    a.tau_eff = coll->tau_eff;
    a.mlp = coll->mlp;
    if (coll->tau_mode < TLBM_TAU_CONST || coll->tau_mode > TLBM_TAU_MLP_BF16)
      return fail(TLBM_ERR_INVALID, "unknown tau_mode %d", coll->tau_mode);
    if ((coll->tau_mode == TLBM_TAU_FIELD || coll->tau_mode == TLBM_TAU_FIELD64) && coll->tau_eff == nullptr)
      return fail(TLBM_ERR_INVALID, "a per-cell tau mode needs tau_eff");
    if (coll->tau_mode == TLBM_TAU_MLP || coll->tau_mode == TLBM_TAU_MLP_BF16) {
      if (coll->mlp == nullptr) return fail(TLBM_ERR_INVALID, "the learned closure needs its packed blob (coll->mlp)");
      if (lattice != TLBM_D3Q19 || coll->kind != TLBM_COLL_BGK)
        return fail(TLBM_ERR_UNSUPPORTED, "the learned closure is built for D3Q19 BGK");
    }
  }
  a.bc_kind = TLBM_BC_NONE;
  a.bc_faces = 0;
  if (bc != nullptr && bc->kind != TLBM_BC_NONE) {
    if (bc->kind == TLBM_BC_CHANNEL_X) {
      if (nx < 3) return fail(TLBM_ERR_SHAPE, "channel boundary needs nx >= 3");
    } else if (bc->kind == TLBM_BC_LID_Y) {
      if (lattice != TLBM_D3Q19) return fail(TLBM_ERR_UNSUPPORTED, "moving lid is D3Q19 only");
    } else if (bc->kind == TLBM_BC_FAR_FIELD) {
      if (lattice == TLBM_D2Q9) return fail(TLBM_ERR_UNSUPPORTED, "far-field faces are 3-D only");
      if (nx < 3) return fail(TLBM_ERR_SHAPE, "far-field boundary needs nx >= 3");
      if (bc->faces & ~63) return fail(TLBM_ERR_INVALID, "unknown far-field face bits 0x%x", bc->faces);
    } else {
      return fail(TLBM_ERR_INVALID, "unknown boundary kind %d", bc->kind);
    }
    a.bc_kind = bc->kind;
    a.bc_faces = bc->faces;
    for (int i = 0; i < 3; ++i) a.bc_u[i] = C(bc->u[i]);
  }
  return TLBM_OK;
}

// ---- instantiation table ----------------------------------------------------

template <class L, int S, int COLL, bool STREAM, int FEAT, int BCK = BCK_RUNTIME>
int launch_lbm(const KArgs<typename Store<S>::C>& a, int bx, int by, uint32_t nb, hipStream_t s) {
  if constexpr ((FEAT & FEAT_MLP) != 0) {
    // nb = tiles_x * tiles_y * nz; the closure step marches TLBM_MLP_ZPER planes per workgroup
    const int zper = mlp_marches<S>() ? TLBM_MLP_ZPER : 1;
    const uint32_t nbz = uint32_t(a.tiles_x) * uint32_t(a.tiles_y) * uint32_t((a.nz + zper - 1) / zper);
    (void)nb;
    hipLaunchKernelGGL((k_lbm_mlp<L, S, COLL, STREAM, FEAT, BCK>), dim3(nbz), dim3(bx, by), 0, s, a, zper);
  } else {
    hipLaunchKernelGGL((k_lbm<L, S, COLL, STREAM, FEAT, BCK>), dim3(nb), dim3(bx, by), 0, s, a);
  }
  return check_launch("k_lbm");
}

// D3Q19 BGK fused steps without a force buffer (the benchmarked and closure
// forms) get one kernel per boundary kind: the uniform state of all kinds and
// of the force sample together spills scalar registers into VGPR lanes, which
// costs the instruction-bound 16-bit and closure forms directly.
template <class L, int S, int COLL, int FEAT>
int launch_lbm_kind(const KArgs<typename Store<S>::C>& a, int bx, int by, uint32_t nb, hipStream_t s) {
  switch (a.bc_kind) {
    case TLBM_BC_CHANNEL_X:
      return launch_lbm<L, S, COLL, true, FEAT, TLBM_BC_CHANNEL_X>(a, bx, by, nb, s);
    case TLBM_BC_LID_Y:
      return launch_lbm<L, S, COLL, true, FEAT, TLBM_BC_LID_Y>(a, bx, by, nb, s);
    case TLBM_BC_FAR_FIELD:
      return launch_lbm<L, S, COLL, true, FEAT, TLBM_BC_FAR_FIELD>(a, bx, by, nb, s);
    default:
      return launch_lbm<L, S, COLL, true, FEAT, TLBM_BC_NONE>(a, bx, by, nb, s);
  }
}

// ---- two steps per launch ----------------------------------------------------

// Tile of the two-step kernel: 64 x 8 output cells (11 waves, 1.29x halo reads;
// D3Q19 fp32: 150 KB for three planes) where three planes fit the 160 KB of LDS
// next to the mask and f_eq tables, else 32 x 8 (6 waves; D3Q27 with 4-byte
// storage).  Measured on the 256^3 channel: 64 x 6 46.0, 64 x 7 50.7, 64 x 8
// 57.2 GLUPS (more waves per CU and less recomputed halo).  32 x 16 recomputes
// less halo (1.20x against 1.29x) and pads uneven grids less, yet ran slower:
// 61.6 against 64.9 GLUPS on the channel and 34.6 against 42.2 on 64 cavities
// of edges 32-160 -- rows of 64 keep each pulled row in whole cache lines.
template <class L, int S>
struct Tile2 {
  static constexpr bool wide = 3 * L::Q * 66 * 10 * sizeof(typename Store<S>::T) <= 158 * 1024;
  static constexpr int TX = wide ? 64 : 32;
  static constexpr int TY = 8;
};

inline int tile2_width(int lattice, int store) {
  return lattice == TLBM_D3Q27 && store == TLBM_STORE_F32 ? 32 : 64;
}

template <class L, int S>
constexpr bool step2_store() {
  return L::Q != 9 && (S == ST_F32 || S == ST_F16_DEV || S == ST_BF16_DEV);
}

// Compute units of the current device (cached per device).
inline int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cache[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// z planes per workgroup.  A two-step workgroup holds most of a CU's LDS, so a CU
// runs one at a time and a launch of W = cols * ceil(nz / zc) workgroups takes
// ceil(W / CUs) rounds of about zc + 2 plane times (zc planes, the two-plane fill
// of the march at half cost, the prologue).  Pick the zc with the least total:
// on MI355X at 256^3 that is 128 (one round of 256 workgroups; 60.7 GLUPS on the
// channel + sphere against 59.1 / 57.2 / 54.8 for zc = 64 / 32 / 16, whose
// costs this model ranks and spaces the same way).
inline int step2_zc(int cols, int nz, int cus) {
  int best = nz;
  int64_t best_cost = INT64_MAX;
  for (int zc = 1; zc <= nz; ++zc) {
    const int64_t w = int64_t(cols) * ((nz + zc - 1) / zc);
    const int64_t cost = (w + cus - 1) / cus * (zc + 2);
    if (cost < best_cost) {
      best_cost = cost;
      best = zc;
    }
  }
  return best;
}

// Batched lid-driven cavities, two steps per launch: the two-step march over the
// column tiles of every cavity in one grid.  Cavity i (edge E, [z][q][y][x] at
// element `offset` with the (y, x) plane padded to a multiple of 64) owns
// workgroups tiles_before .. tiles_before + ceil(E/TX) ceil(E/TY) ceil(E/zc) - 1;
// a workgroup finds its cavity by a binary search over those prefix sums and
// runs the single-grid march with the cavity's strides, relaxation rate, the
// Zou-He lid and box walls -- per cavity what tlbm_step2 computes for that
// cavity alone with its wall mask and TLBM_BC_LID_Y.
struct Cav2Dev {
  int64_t offset;
  uint32_t tiles_before;
  int32_t edge;
  double omega;
};

template <typename C>
struct Cav2Args {
  KArgs<C> k;  // fields shared by every cavity (buffers, lid, collision mode)
  uint32_t total;
  int n, zc;
  Cav2Dev c[TLBM_CAVITIES_MAX];
};
static_assert(sizeof(Cav2Args<float>) < 4000, "cavity table travels in the kernel arguments");

template <int S>
__global__ __launch_bounds__((tb_threads<Tile2<D3Q19, S>::TX, Tile2<D3Q19, S>::TY>())) void k_cav2(
    const Cav2Args<typename Store<S>::C> a) {
  using T = typename Store<S>::T;
  using C = typename Store<S>::C;
  const uint32_t g = xcd_tile(blockIdx.x, gridDim.x);
  int lo = 0, hi = a.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.c[mid].tiles_before <= g) lo = mid; else hi = mid - 1;
  }
  const int E = a.c[lo].edge;
  const int64_t plane = (int64_t(E) * E + 63) / 64 * 64;
  KArgs<C> k = a.k;
  k.fin = static_cast<const T*>(a.k.fin) + a.c[lo].offset;
  k.fout = static_cast<T*>(a.k.fout) + a.c[lo].offset;
  k.nx = k.ny = k.nz = E;
  k.isq = k.osq = plane;
  k.isz = k.osz = D3Q19::Q * plane;
  k.isy = k.osy = E;
  k.omega = C(a.c[lo].omega);
  k.tau = C(1.0 / a.c[lo].omega);
  lbm2_body<D3Q19, S, TLBM_COLL_BGK, FEAT_BC, TLBM_BC_LID_Y, Tile2<D3Q19, S>::TX, Tile2<D3Q19, S>::TY, true>(
      k, a.zc, g - a.c[lo].tiles_before);
}

// Workgroups of a two-step cavity launch with z chunks of zc planes, per cavity
// prefix sums into `a`; -1 when the count would not fit 31 bits.
template <typename C, int TX, int TY>
int64_t cav2_layout(Cav2Args<C>& a, const tlbm_cavity* cavs, int n, int zc) {
  uint64_t w = 0;
  for (int i = 0; i < n; ++i) {
    const int E = int(cavs[i].edge);
    a.c[i].offset = cavs[i].offset;
    a.c[i].tiles_before = uint32_t(w);
    a.c[i].edge = E;
    a.c[i].omega = 1.0 / cavs[i].tau;
    w += uint64_t((E + TX - 1) / TX) * uint64_t((E + TY - 1) / TY) * uint64_t((E + zc - 1) / zc);
    if (w >= (uint64_t(1) << 31)) return -1;
  }
  return int64_t(w);
}

// z planes per workgroup for a batch of cavities: the rounds x planes model of
// step2_zc over the workgroups of all cavities together
template <int TX, int TY>
int cav2_zc(const tlbm_cavity* cavs, int n, int cus) {
  int emax = 1;
  for (int i = 0; i < n; ++i) emax = cavs[i].edge > emax ? int(cavs[i].edge) : emax;
  int best = emax;
  int64_t best_cost = INT64_MAX;
  for (int zc = 1; zc <= emax; ++zc) {
    int64_t w = 0;
    for (int i = 0; i < n; ++i) {
      const int E = int(cavs[i].edge);
      w += int64_t((E + TX - 1) / TX) * ((E + TY - 1) / TY) * ((E + zc - 1) / zc);
    }
    const int64_t cost = (w + cus - 1) / cus * (zc + 2);
    if (cost < best_cost) {
      best_cost = cost;
      best = zc;
    }
  }
  return best;
}

template <int S>
int launch_cav2(const void* fin, void* fout, const tlbm_cavity* cavs, int n, double u_lid, int zc, hipStream_t s) {
  using C = typename Store<S>::C;
  constexpr int TX = Tile2<D3Q19, S>::TX, TY = Tile2<D3Q19, S>::TY;
  Cav2Args<C> a;
  memset(&a, 0, sizeof(a));
  a.k.fin = fin;
  a.k.fout = fout;
  a.k.tau_mode = TLBM_TAU_CONST;
  a.k.eq_floor = 1;
  a.k.bc_kind = TLBM_BC_LID_Y;
  a.k.bc_u[0] = C(u_lid);
  a.n = n;
  a.zc = zc > 0 ? zc : cav2_zc<TX, TY>(cavs, n, device_cus());
  const int64_t w = cav2_layout<C, TX, TY>(a, cavs, n, a.zc);
  if (w < 0) return fail(TLBM_ERR_SHAPE, "too many workgroups for one launch");
  a.total = uint32_t(w);
  hipLaunchKernelGGL(k_cav2<S>, dim3(a.total), dim3(tb_threads<TX, TY>()), 0, s, a);
  return check_launch("k_cav2");
}

// why a two-step launch cannot serve these arguments (nullptr: it can)
inline const char* step2_reason(int lattice, int store, const tlbm_collision* coll, const tlbm_boundary* bc, int nx) {
  if (lattice == TLBM_D2Q9) return "two-step launches are built for D3Q19 and D3Q27";
  if (store != TLBM_STORE_F32 && store != TLBM_STORE_F16_DEV && store != TLBM_STORE_BF16_DEV)
    return "two-step launches take float32 or 16-bit deviation storage";
  if (coll == nullptr || coll->kind == TLBM_COLL_NONE) return "collision parameters required";
  if (coll->tau_mode == TLBM_TAU_MLP || coll->tau_mode == TLBM_TAU_MLP_BF16)
    return "the learned closure runs in the one-step kernel";
  if (coll->tau_mode == TLBM_TAU_FIELD64) return "the fused step takes a float32 tau_eff field";
  if (bc != nullptr && (bc->kind == TLBM_BC_CHANNEL_X || bc->kind == TLBM_BC_FAR_FIELD)) {
    // the outlet copies x = nx-2, whose pulls reach x = nx-3: same tile as x = nx-1
    if (nx < 4 || (nx - 1) % tile2_width(lattice, store) < 2) return "the outlet column would straddle two tiles";
  }
  return nullptr;
}

template <class L, int S, int COLL, int FEAT, int BCK>
int launch_lbm2(const KArgs<typename Store<S>::C>& a, hipStream_t s) {
  constexpr int TX = Tile2<L, S>::TX, TY = Tile2<L, S>::TY;
  const int gx = (a.nx + TX - 1) / TX, gy = (a.ny + TY - 1) / TY;
  const int zc = step2_zc(gx * gy, a.nz, device_cus());
  const uint32_t tiles = uint32_t(gx) * uint32_t(gy) * uint32_t((a.nz + zc - 1) / zc);
  hipLaunchKernelGGL((k_lbm2<L, S, COLL, FEAT, BCK, TX, TY>), dim3(tiles), dim3(tb_threads<TX, TY>()), 0, s, a, zc);
  return check_launch("k_lbm2");
}

// D3Q19 builds one two-step kernel per boundary kind: each kind's uniform state
// then fits the scalar registers (all kinds in one kernel spilled ~60 SGPRs into
// VGPR lanes, doubling the VALU work of the z loop); D3Q27 reads the kind at run time.
template <class L, int S, int COLL>
int launch_lbm2_bc(const KArgs<typename Store<S>::C>& a, hipStream_t s) {
  if constexpr (L::Q == 19) {
    switch (a.bc_kind) {
      case TLBM_BC_CHANNEL_X:
        return launch_lbm2<L, S, COLL, FEAT_BC, TLBM_BC_CHANNEL_X>(a, s);
      case TLBM_BC_LID_Y:
        return launch_lbm2<L, S, COLL, FEAT_BC, TLBM_BC_LID_Y>(a, s);
      case TLBM_BC_FAR_FIELD:
        return launch_lbm2<L, S, COLL, FEAT_BC, TLBM_BC_FAR_FIELD>(a, s);
      default:
        return launch_lbm2<L, S, COLL, FEAT_BC, TLBM_BC_NONE>(a, s);
    }
  } else {
    return launch_lbm2<L, S, COLL, FEAT_BC, BCK_RUNTIME>(a, s);
  }
}

template <class L, int COLL>
constexpr bool coll_built() {
  if (COLL == TLBM_COLL_NONE || COLL == TLBM_COLL_BGK || COLL == TLBM_COLL_MRT) return true;
  if (COLL == TLBM_COLL_CM) return true;
  if (COLL == TLBM_COLL_CUMULANT) return L::Q != 9;
  if (COLL == TLBM_COLL_GEIER) return L::Q == 27;
  return false;
}

struct Call {
  int lattice, store, kind;
  bool stream_, feat;
  int steps;
  const void* fin;
  const int64_t* is;
  void* fout;
  const int64_t* os;
  int nz, ny, nx;
  const tlbm_collision* coll;
  const int8_t* solid;
  const tlbm_boundary* bc;
  double* force;
  hipStream_t s;
};

template <class L, int S, int COLL>
int run_coll(const Call& c) {
  if constexpr (!coll_built<L, COLL>()) {
    return fail(TLBM_ERR_UNSUPPORTED, "collision %d is not built for Q=%d", COLL, L::Q);
  } else {
    KArgs<typename Store<S>::C> a;
    int bx, by;
    uint32_t nb;
    int rc = fill_args(a, c.lattice, c.fin, c.is, c.fout, c.os, c.nz, c.ny, c.nx, c.coll, c.solid, c.bc, c.force, &bx,
                       &by, &nb);
    if (rc != TLBM_OK) return rc;
    if (c.coll != nullptr && c.coll->tau_mode == TLBM_TAU_MLP) {
      if constexpr (L::Q == 19 && COLL == TLBM_COLL_BGK) {
        if (!c.stream_) return fail(TLBM_ERR_UNSUPPORTED, "the learned closure runs in the fused step");
        if (c.feat && c.force == nullptr) return launch_lbm_kind<L, S, COLL, FEAT_MLP | FEAT_BC>(a, bx, by, nb, c.s);
        if (c.feat) return launch_lbm<L, S, COLL, true, FEAT_MLP_ALL>(a, bx, by, nb, c.s);
        return launch_lbm<L, S, COLL, true, FEAT_MLP>(a, bx, by, nb, c.s);
      } else {
        return fail(TLBM_ERR_UNSUPPORTED, "the learned closure is built for D3Q19 BGK");
      }
    }
    if (c.coll != nullptr && c.coll->tau_mode == TLBM_TAU_MLP_BF16) {
      // bf16-operand closure: fp32, fp64 and 16-bit deviation storage
      if constexpr (L::Q == 19 && COLL == TLBM_COLL_BGK &&
                    (S == ST_F32 || S == ST_F64 || S == ST_F16_DEV || S == ST_BF16_DEV)) {
        constexpr int M = FEAT_MLP | FEAT_MLP_BF16;
        if (!c.stream_) return fail(TLBM_ERR_UNSUPPORTED, "the learned closure runs in the fused step");
        if (c.feat && c.force == nullptr) return launch_lbm_kind<L, S, COLL, M | FEAT_BC>(a, bx, by, nb, c.s);
        if (c.feat) return launch_lbm<L, S, COLL, true, M | FEAT_ALL>(a, bx, by, nb, c.s);
        return launch_lbm<L, S, COLL, true, M>(a, bx, by, nb, c.s);
      } else {
        return fail(TLBM_ERR_UNSUPPORTED,
                    "the bf16 closure is built for D3Q19 BGK with fp32, fp64 or 16-bit deviation storage");
      }
    }
    if (c.steps == 2) {
      if constexpr (step2_store<L, S>() && COLL != TLBM_COLL_NONE) {
        if (c.force != nullptr) {
          // the Ladd force sample (a D3Q19 obstacle quantity, triton_fused_obstacle.py:1494-1512):
          // the channel gets its own BGK kernel, the rest read the boundary kind at run time
          if constexpr (L::Q == 19) {
            if constexpr (COLL == TLBM_COLL_BGK) {
              if (a.bc_kind == TLBM_BC_CHANNEL_X) return launch_lbm2<L, S, COLL, FEAT_ALL, TLBM_BC_CHANNEL_X>(a, c.s);
            }
            return launch_lbm2<L, S, COLL, FEAT_ALL, BCK_RUNTIME>(a, c.s);
          } else {
            return fail(TLBM_ERR_UNSUPPORTED, "the two-step force sample is built for D3Q19");
          }
        }
        if (c.feat) return launch_lbm2_bc<L, S, COLL>(a, c.s);
        return launch_lbm2<L, S, COLL, FEAT_NONE, TLBM_BC_NONE>(a, c.s);
      } else {
        return fail(TLBM_ERR_UNSUPPORTED, "two-step launch not built for this lattice / storage");
      }
    }
    if (c.stream_) {
      if constexpr (L::Q == 19 && COLL == TLBM_COLL_BGK) {
        if (c.feat && c.force == nullptr) return launch_lbm_kind<L, S, COLL, FEAT_BC>(a, bx, by, nb, c.s);
      }
      if (c.feat) return launch_lbm<L, S, COLL, true, FEAT_ALL>(a, bx, by, nb, c.s);
      return launch_lbm<L, S, COLL, true, FEAT_NONE>(a, bx, by, nb, c.s);
    }
    if constexpr (S == ST_F32 || S == ST_F64) {
      return launch_lbm<L, S, COLL, false, FEAT_NONE>(a, bx, by, nb, c.s);
    } else {
      return fail(TLBM_ERR_DTYPE, "collide-only ops take float32 or float64 populations");
    }
  }
}

template <class L, int S>
int run_store(const Call& c) {
  switch (c.kind) {
    case TLBM_COLL_NONE:
      return run_coll<L, S, TLBM_COLL_NONE>(c);
    case TLBM_COLL_BGK:
      return run_coll<L, S, TLBM_COLL_BGK>(c);
    case TLBM_COLL_MRT:
      return run_coll<L, S, TLBM_COLL_MRT>(c);
    case TLBM_COLL_CM:
      return run_coll<L, S, TLBM_COLL_CM>(c);
    case TLBM_COLL_CUMULANT:
      return run_coll<L, S, TLBM_COLL_CUMULANT>(c);
    case TLBM_COLL_GEIER:
      return run_coll<L, S, TLBM_COLL_GEIER>(c);
    default:
      return fail(TLBM_ERR_INVALID, "unknown collision kind %d", c.kind);
  }
}

template <class L>
int run_lattice(const Call& c) {
  switch (c.store) {
    case ST_F32:
      return run_store<L, ST_F32>(c);
    case ST_F64:
      return run_store<L, ST_F64>(c);
    case ST_F16:
      return run_store<L, ST_F16>(c);
    case ST_BF16:
      return run_store<L, ST_BF16>(c);
    case ST_F16_DEV:
      return run_store<L, ST_F16_DEV>(c);
    case ST_BF16_DEV:
      return run_store<L, ST_BF16_DEV>(c);
    default:
      return fail(TLBM_ERR_DTYPE, "unknown storage %d", c.store);
  }
}

int run(const Call& c) {
  switch (c.lattice) {
    case TLBM_D2Q9:
      return run_lattice<D2Q9>(c);
    case TLBM_D3Q19:
      return run_lattice<D3Q19>(c);
    case TLBM_D3Q27:
      return run_lattice<D3Q27>(c);
    default:
      return fail(TLBM_ERR_INVALID, "unknown lattice Q=%d", c.lattice);
  }
}

template <class F>
int by_lattice(int lattice, F&& fn) {
  switch (lattice) {
    case TLBM_D2Q9:
      return fn(D2Q9{});
    case TLBM_D3Q19:
      return fn(D3Q19{});
    case TLBM_D3Q27:
      return fn(D3Q27{});
    default:
      return fail(TLBM_ERR_INVALID, "unknown lattice Q=%d", lattice);
  }
}

template <class L, class F>
int by_store(int store, F&& fn) {
  switch (store) {
    case ST_F32:
      return fn(std::integral_constant<int, ST_F32>{});
    case ST_F64:
      return fn(std::integral_constant<int, ST_F64>{});
    case ST_F16:
      return fn(std::integral_constant<int, ST_F16>{});
    case ST_BF16:
      return fn(std::integral_constant<int, ST_BF16>{});
    case ST_F16_DEV:
      return fn(std::integral_constant<int, ST_F16_DEV>{});
    case ST_BF16_DEV:
      return fn(std::integral_constant<int, ST_BF16_DEV>{});
    default:
      return fail(TLBM_ERR_DTYPE, "unknown storage %d", store);
  }
}

uint32_t blocks_for(uint64_t n) { return uint32_t((n + 255) / 256); }

int prepare(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
            const int64_t* out_strides, int nz, int ny, int nx, int64_t is[3], int64_t os[3]) {
  int rc = check_dims(lattice, nz, ny, nx);
  if (rc) return rc;
  if (!store_ok(store)) return fail(TLBM_ERR_DTYPE, "unknown storage %d", store);
  if (f_in == nullptr || f_out == nullptr) return fail(TLBM_ERR_INVALID, "null population pointer");
  rc = resolve_strides(in_strides, lattice, nz, ny, nx, store_size(store), is, "f_in");
  if (rc) return rc;
  return resolve_strides(out_strides, lattice, nz, ny, nx, store_size(store), os, "f_out");
}

}  // namespace

namespace tlbm {
// two steps of every cavity of a validated batch (cavities.hip checks the table)
int cav2_step(int store, const void* fin, void* fout, const tlbm_cavity* cavs, int n, double u_lid, int zc,
              hipStream_t s) {
  switch (store) {
    case TLBM_STORE_F32: return launch_cav2<ST_F32>(fin, fout, cavs, n, u_lid, zc, s);
    case TLBM_STORE_F16_DEV: return launch_cav2<ST_F16_DEV>(fin, fout, cavs, n, u_lid, zc, s);
    case TLBM_STORE_BF16_DEV: return launch_cav2<ST_BF16_DEV>(fin, fout, cavs, n, u_lid, zc, s);
    default: return fail(TLBM_ERR_UNSUPPORTED, "two steps per launch take float32 or 16-bit deviation storage");
  }
}
}  // namespace tlbm

extern "C" {

int tlbm_version(void) { return 2; }

const char* tlbm_last_error(void) { return g_err; }

const char* tlbm_status_string(int status) {
  switch (status) {
    case TLBM_OK:
      return "ok";
    case TLBM_ERR_INVALID:
      return "invalid argument";
    case TLBM_ERR_SHAPE:
      return "shape error";
    case TLBM_ERR_DTYPE:
      return "unsupported dtype";
    case TLBM_ERR_UNSUPPORTED:
      return "unsupported combination";
    case TLBM_ERR_ALIAS:
      return "input and output alias";
    case TLBM_ERR_HIP:
      return "HIP runtime error";
    default:
      return "unknown status";
  }
}

int tlbm_step(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
              const int64_t* out_strides, int nz, int ny, int nx, const tlbm_collision* coll, const int8_t* solid,
              const tlbm_boundary* bc, double* force, tlbm_stream_t stream) {
  g_err[0] = 0;
  int64_t is[3], os[3];
  int rc = prepare(lattice, store, f_in, in_strides, f_out, out_strides, nz, ny, nx, is, os);
  if (rc) return rc;
  if (f_in == f_out) return fail(TLBM_ERR_ALIAS, "f_in and f_out alias; a pull step needs two lattices");
  if (coll == nullptr) return fail(TLBM_ERR_INVALID, "collision parameters required");
  if (coll->tau_mode == TLBM_TAU_FIELD64)
    return fail(TLBM_ERR_UNSUPPORTED, "the fused step takes a float32 tau_eff field (TLBM_TAU_FIELD)");
  const bool feat = solid != nullptr || force != nullptr || (bc != nullptr && bc->kind != TLBM_BC_NONE);
  if (force != nullptr && solid == nullptr) return fail(TLBM_ERR_INVALID, "force needs a solid mask");
  Call c{lattice, store, coll->kind, true, feat, 1, f_in, is, f_out, os, nz, ny, nx, coll, solid, bc, force,
         (hipStream_t)stream};
  return run(c);
}

int tlbm_step2(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
               const int64_t* out_strides, int nz, int ny, int nx, const tlbm_collision* coll, const int8_t* solid,
               const tlbm_boundary* bc, double* force, tlbm_stream_t stream) {
  g_err[0] = 0;
  int64_t is[3], os[3];
  int rc = prepare(lattice, store, f_in, in_strides, f_out, out_strides, nz, ny, nx, is, os);
  if (rc) return rc;
  if (f_in == f_out) return fail(TLBM_ERR_ALIAS, "f_in and f_out alias; a pull step needs two lattices");
  if (coll == nullptr) return fail(TLBM_ERR_INVALID, "collision parameters required");
  if (const char* why = step2_reason(lattice, store, coll, bc, nx)) return fail(TLBM_ERR_UNSUPPORTED, "%s", why);
  if (force != nullptr && solid == nullptr) return fail(TLBM_ERR_INVALID, "force needs a solid mask");
  if (force != nullptr && lattice != TLBM_D3Q19)
    return fail(TLBM_ERR_UNSUPPORTED, "the two-step force sample is built for D3Q19; sample D3Q27 with tlbm_step");
  const bool feat = solid != nullptr || force != nullptr || (bc != nullptr && bc->kind != TLBM_BC_NONE);
  Call c{lattice, store, coll->kind, true, feat, 2, f_in, is, f_out, os, nz, ny, nx, coll, solid, bc, force,
         (hipStream_t)stream};
  return run(c);
}

int tlbm_step2_supported(int lattice, int store, const tlbm_collision* coll, const tlbm_boundary* bc, int nx) {
  return step2_reason(lattice, store, coll, bc, nx) == nullptr ? 1 : 0;
}

int tlbm_stream(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
                const int64_t* out_strides, int nz, int ny, int nx, tlbm_stream_t stream) {
  g_err[0] = 0;
  int64_t is[3], os[3];
  int rc = prepare(lattice, store, f_in, in_strides, f_out, out_strides, nz, ny, nx, is, os);
  if (rc) return rc;
  if (f_in == f_out) return fail(TLBM_ERR_ALIAS, "f_in and f_out alias; streaming needs two buffers");
  Call c{lattice, store, TLBM_COLL_NONE, true, false, 1, f_in, is, f_out, os, nz, ny, nx, nullptr, nullptr, nullptr,
         nullptr, (hipStream_t)stream};
  return run(c);
}

int tlbm_collide(int lattice, int store, const void* f_in, const int64_t* in_strides, void* f_out,
                 const int64_t* out_strides, int nz, int ny, int nx, const tlbm_collision* coll,
                 tlbm_stream_t stream) {
  g_err[0] = 0;
  if (store != TLBM_STORE_F32 && store != TLBM_STORE_F64)
    return fail(TLBM_ERR_DTYPE, "collide-only ops take float32 or float64 populations");
  int64_t is[3], os[3];
  int rc = prepare(lattice, store, f_in, in_strides, f_out, out_strides, nz, ny, nx, is, os);
  if (rc) return rc;
  if (coll == nullptr || coll->kind == TLBM_COLL_NONE) return fail(TLBM_ERR_INVALID, "collision kind required");
  Call c{lattice, store, coll->kind, false, false, 1, f_in, is, f_out, os, nz, ny, nx, coll, nullptr, nullptr, nullptr,
         (hipStream_t)stream};
  return run(c);
}

int tlbm_equilibrium(int lattice, int dtype, const void* rho, const void* ux, const void* uy, const void* uz,
                     void* f_out, int64_t ncells, tlbm_stream_t stream) {
  g_err[0] = 0;
  if (ncells < 0) return fail(TLBM_ERR_SHAPE, "negative cell count");
  if (ncells == 0) return TLBM_OK;
  if (!rho || !ux || !uy || !f_out) return fail(TLBM_ERR_INVALID, "null pointer");
  if (lattice != TLBM_D2Q9 && uz == nullptr) return fail(TLBM_ERR_INVALID, "uz required for 3-D lattices");
  hipStream_t s = (hipStream_t)stream;
  const uint64_t n = uint64_t(ncells);
  return by_lattice(lattice, [&](auto lat) -> int {
    using L = decltype(lat);
    if (dtype == TLBM_STORE_F32) {
      hipLaunchKernelGGL((k_equilibrium<L, float>), dim3(blocks_for(n)), dim3(256), 0, s, (const float*)rho,
                         (const float*)ux, (const float*)uy, (const float*)uz, (float*)f_out, n);
    } else if (dtype == TLBM_STORE_F64) {
      hipLaunchKernelGGL((k_equilibrium<L, double>), dim3(blocks_for(n)), dim3(256), 0, s, (const double*)rho,
                         (const double*)ux, (const double*)uy, (const double*)uz, (double*)f_out, n);
    } else {
      return fail(TLBM_ERR_DTYPE, "equilibrium takes float32 or float64 fields");
    }
    return check_launch("k_equilibrium");
  });
}

int tlbm_macroscopic(int lattice, int dtype, const void* f, void* rho, void* ux, void* uy, void* uz,
                     int64_t ncells, tlbm_stream_t stream) {
  g_err[0] = 0;
  if (ncells < 0) return fail(TLBM_ERR_SHAPE, "negative cell count");
  if (ncells == 0) return TLBM_OK;
  if (!f || !rho || !ux || !uy) return fail(TLBM_ERR_INVALID, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  const uint64_t n = uint64_t(ncells);
  return by_lattice(lattice, [&](auto lat) -> int {
    using L = decltype(lat);
    if (dtype == TLBM_STORE_F32) {
      hipLaunchKernelGGL((k_macroscopic<L, float>), dim3(blocks_for(n)), dim3(256), 0, s, (const float*)f,
                         (float*)rho, (float*)ux, (float*)uy, (float*)uz, n);
    } else if (dtype == TLBM_STORE_F64) {
      hipLaunchKernelGGL((k_macroscopic<L, double>), dim3(blocks_for(n)), dim3(256), 0, s, (const double*)f,
                         (double*)rho, (double*)ux, (double*)uy, (double*)uz, n);
    } else {
      return fail(TLBM_ERR_DTYPE, "macroscopic takes float32 or float64 populations");
    }
    return check_launch("k_macroscopic");
  });
}

int tlbm_bounce_back(int lattice, int dtype, const void* f_in, void* f_out, const int8_t* mask, int64_t ncells,
                     tlbm_stream_t stream) {
  g_err[0] = 0;
  if (ncells < 0) return fail(TLBM_ERR_SHAPE, "negative cell count");
  if (ncells == 0) return TLBM_OK;
  if (!f_in || !f_out || !mask) return fail(TLBM_ERR_INVALID, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  const uint64_t n = uint64_t(ncells);
  return by_lattice(lattice, [&](auto lat) -> int {
    using L = decltype(lat);
    if (dtype == TLBM_STORE_F32) {
      hipLaunchKernelGGL((k_bounce_back<L, float>), dim3(blocks_for(n)), dim3(256), 0, s, (const float*)f_in,
                         (float*)f_out, mask, n);
    } else if (dtype == TLBM_STORE_F64) {
      hipLaunchKernelGGL((k_bounce_back<L, double>), dim3(blocks_for(n)), dim3(256), 0, s, (const double*)f_in,
                         (double*)f_out, mask, n);
    } else {
      return fail(TLBM_ERR_DTYPE, "bounce_back takes float32 or float64 populations");
    }
    return check_launch("k_bounce_back");
  });
}

static int launch_sums(int lattice, int store, const void* f, const int64_t* strides, int nz, int ny, int nx,
                       const int8_t* solid, double* out, hipStream_t s) {
  int rc = check_dims(lattice, nz, ny, nx);
  if (rc) return rc;
  if (!store_ok(store)) return fail(TLBM_ERR_DTYPE, "unknown storage %d", store);
  if (!f || !out) return fail(TLBM_ERR_INVALID, "null pointer");
  int64_t st[3];
  rc = resolve_strides(strides, lattice, nz, ny, nx, store_size(store), st, "f");
  if (rc) return rc;
  const Layout lay{st[0], st[1], st[2], nx, ny, nz};
  const uint64_t n = uint64_t(nz) * ny * nx;
  uint32_t nb = blocks_for(n);
  if (nb > 4096) nb = 4096;
  return by_lattice(lattice, [&](auto lat) -> int {
    using L = decltype(lat);
    return by_store<L>(store, [&](auto sk) -> int {
      constexpr int S = decltype(sk)::value;
      hipLaunchKernelGGL((k_moment_sums<L, S>), dim3(nb), dim3(256), 0, s, f, lay, solid, out);
      return check_launch("k_moment_sums");
    });
  });
}

int tlbm_moment_sums(int lattice, int store, const void* f, const int64_t* strides, int nz, int ny, int nx,
                     double* out, tlbm_stream_t stream) {
  g_err[0] = 0;
  return launch_sums(lattice, store, f, strides, nz, ny, nx, nullptr, out, (hipStream_t)stream);
}

int tlbm_obstacle_force(int lattice, int store, const void* f, const int64_t* strides, int nz, int ny, int nx,
                        const int8_t* solid, double* out, tlbm_stream_t stream) {
  g_err[0] = 0;
  if (!solid) return fail(TLBM_ERR_INVALID, "null solid mask");
  return launch_sums(lattice, store, f, strides, nz, ny, nx, solid, out, (hipStream_t)stream);
}

int tlbm_smagorinsky_tau_sum(int lattice, int store, const void* f, const int64_t* strides, int nz, int ny, int nx,
                             double tau, double cs_delta_sq, double* out, tlbm_stream_t stream) {
  g_err[0] = 0;
  int rc = check_dims(lattice, nz, ny, nx);
  if (rc) return rc;
  if (!store_ok(store)) return fail(TLBM_ERR_DTYPE, "unknown storage %d", store);
  if (!f || !out) return fail(TLBM_ERR_INVALID, "null pointer");
  if (!(tau > 0.5)) return fail(TLBM_ERR_INVALID, "tau must be > 0.5 (got %g)", tau);
  int64_t st[3];
  rc = resolve_strides(strides, lattice, nz, ny, nx, store_size(store), st, "f");
  if (rc) return rc;
  const Layout lay{st[0], st[1], st[2], nx, ny, nz};
  uint32_t nb = blocks_for(uint64_t(nz) * ny * nx);
  if (nb > 4096) nb = 4096;
  hipStream_t s = (hipStream_t)stream;
  return by_lattice(lattice, [&](auto lat) -> int {
    using L = decltype(lat);
    return by_store<L>(store, [&](auto sk) -> int {
      constexpr int S = decltype(sk)::value;
      hipLaunchKernelGGL((k_tau_sum<L, S>), dim3(nb), dim3(256), 0, s, f, lay, tau, cs_delta_sq, out);
      return check_launch("k_tau_sum");
    });
  });
}

int tlbm_scale(int lattice, int store, void* f, const int64_t* strides, int nz, int ny, int nx, const double* scale,
               tlbm_stream_t stream) {
  g_err[0] = 0;
  int rc = check_dims(lattice, nz, ny, nx);
  if (rc) return rc;
  if (!store_ok(store)) return fail(TLBM_ERR_DTYPE, "unknown storage %d", store);
  if (!f || !scale) return fail(TLBM_ERR_INVALID, "null pointer");
  int64_t st[3];
  rc = resolve_strides(strides, lattice, nz, ny, nx, store_size(store), st, "f");
  if (rc) return rc;
  const Layout lay{st[0], st[1], st[2], nx, ny, nz};
  hipStream_t s = (hipStream_t)stream;
  const uint64_t n = uint64_t(nz) * ny * nx;
  return by_lattice(lattice, [&](auto lat) -> int {
    using L = decltype(lat);
    return by_store<L>(store, [&](auto sk) -> int {
      constexpr int S = decltype(sk)::value;
      hipLaunchKernelGGL((k_scale<L, S>), dim3(blocks_for(n)), dim3(256), 0, s, f, lay, scale);
      return check_launch("k_scale");
    });
  });
}

int tlbm_boundary_eager(int dtype, const void* f_in, void* f_out, int nz, int ny, int nx, const tlbm_boundary* bc,
                        const int8_t* mask_a, const int8_t* mask_b, tlbm_stream_t stream) {
  g_err[0] = 0;
  if (nz < 1 || ny < 1 || nx < 1) return fail(TLBM_ERR_SHAPE, "grid dims must be >= 1 (got %d,%d,%d)", nz, ny, nx);
  const uint64_t n = uint64_t(nz) * uint64_t(ny) * uint64_t(nx);
  if (n >= (uint64_t(1) << 31)) return fail(TLBM_ERR_SHAPE, "grid of %llu cells exceeds 2^31", (unsigned long long)n);
  if (!f_in || !f_out) return fail(TLBM_ERR_INVALID, "null pointer");
  if (f_in == f_out) return fail(TLBM_ERR_ALIAS, "f_in and f_out alias; the boundary op writes a new state");
  const int kind = bc != nullptr ? bc->kind : TLBM_BC_NONE;
  if (kind != TLBM_BC_NONE && kind != TLBM_BC_CHANNEL_X && kind != TLBM_BC_FAR_FIELD && kind != TLBM_BC_LID_Y)
    return fail(TLBM_ERR_INVALID, "unknown boundary kind %d", kind);
  if ((kind == TLBM_BC_CHANNEL_X || kind == TLBM_BC_FAR_FIELD) && nx < 3)
    return fail(TLBM_ERR_SHAPE, "inlet / outlet boundaries need nx >= 3 (got %d)", nx);
  const int faces = bc != nullptr ? bc->faces : 0;
  const double u0 = bc != nullptr ? bc->u[0] : 0.0, u1 = bc != nullptr ? bc->u[1] : 0.0,
               u2 = bc != nullptr ? bc->u[2] : 0.0;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TLBM_STORE_F32) {
    hipLaunchKernelGGL(k_boundary<float>, dim3(blocks_for(n)), dim3(256), 0, s, (const float*)f_in, (float*)f_out, nz,
                       ny, nx, kind, faces, float(u0), float(u1), float(u2), mask_a, mask_b);
  } else if (dtype == TLBM_STORE_F64) {
    hipLaunchKernelGGL(k_boundary<double>, dim3(blocks_for(n)), dim3(256), 0, s, (const double*)f_in,
                       (double*)f_out, nz, ny, nx, kind, faces, u0, u1, u2, mask_a, mask_b);
  } else {
    return fail(TLBM_ERR_DTYPE, "boundary ops take float32 or float64 populations");
  }
  return check_launch("k_boundary");
}

int tlbm_sphere_mask(int nx, int ny, int nz, double cx, double cy, double cz, double radius, uint8_t* out,
                     tlbm_stream_t stream) {
  g_err[0] = 0;
  if (nz < 1 || ny < 1 || nx < 1) return fail(TLBM_ERR_SHAPE, "grid dims must be >= 1 (got %d,%d,%d)", nz, ny, nx);
  if (!out) return fail(TLBM_ERR_INVALID, "null pointer");
  const uint64_t n = uint64_t(nz) * uint64_t(ny) * uint64_t(nx);
  if (n >= (uint64_t(1) << 31)) return fail(TLBM_ERR_SHAPE, "grid of %llu cells exceeds 2^31", (unsigned long long)n);
  hipLaunchKernelGGL(k_sphere_mask, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, float(cx),
                     float(cy), float(cz), float(radius * radius), out);
  return check_launch("k_sphere_mask");
}

int tlbm_gradient_closure_tau(int dtype, const void* ux, const void* uy, const void* uz, const int8_t* solid,
                              const double* wall_velocity, int nz, int ny, int nx, int model, double coef, double tau,
                              void* tau_out, tlbm_stream_t stream) {
  g_err[0] = 0;
  if (nz < 1 || ny < 1 || nx < 1) return fail(TLBM_ERR_SHAPE, "grid dims must be >= 1 (got %d,%d,%d)", nz, ny, nx);
  const uint64_t n = uint64_t(nz) * uint64_t(ny) * uint64_t(nx);
  if (n >= (uint64_t(1) << 31)) return fail(TLBM_ERR_SHAPE, "grid of %llu cells exceeds 2^31", (unsigned long long)n);
  if (!ux || !uy || !uz || !tau_out) return fail(TLBM_ERR_INVALID, "null pointer");
  if (model != TLBM_SGS_WALE && model != TLBM_SGS_VREMAN) return fail(TLBM_ERR_INVALID, "unknown closure %d", model);
  if (!(coef >= 0.0)) return fail(TLBM_ERR_INVALID, "closure coefficient must be non-negative (got %g)", coef);
  if (!(tau > 0.5)) return fail(TLBM_ERR_INVALID, "tau must be > 0.5 (got %g)", tau);
  double w[3] = {0.0, 0.0, 0.0};
  if (wall_velocity != nullptr)
    for (int k = 0; k < 3; ++k) w[k] = wall_velocity[k];
  const double cap = tau + 0.5 > 1.0 ? tau + 0.5 : 1.0;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == TLBM_STORE_F32) {
    hipLaunchKernelGGL((k_gradient_tau<float>), dim3(blocks_for(n)), dim3(256), 0, s, (const float*)ux,
                       (const float*)uy, (const float*)uz, solid, float(w[0]), float(w[1]), float(w[2]), nz, ny, nx,
                       model, float(coef), float(tau), float(cap), (float*)tau_out);
  } else if (dtype == TLBM_STORE_F64) {
    hipLaunchKernelGGL((k_gradient_tau<double>), dim3(blocks_for(n)), dim3(256), 0, s, (const double*)ux,
                       (const double*)uy, (const double*)uz, solid, w[0], w[1], w[2], nz, ny, nx, model, coef, tau,
                       cap, (double*)tau_out);
  } else {
    return fail(TLBM_ERR_DTYPE, "the gradient closures take float32 or float64 velocity fields");
  }
  return check_launch("k_gradient_tau");
}

}  // extern "C"
