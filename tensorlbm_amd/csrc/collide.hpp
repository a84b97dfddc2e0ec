// Register-resident collision operators, templated on the velocity set L and
// the arithmetic type C (float or double).  Each operator maps the Q
// post-stream populations f[] of one cell to their post-collision values in
// place.  Lattice constants and moment tables are compile-time, so unrolled
// loops fold every zero / unit coefficient.
#pragma once

#include "lattice.hpp"
#include "tables.inc"

namespace tlbm {

template <typename C>
struct Macro {
  C rho, rho_safe, ux, uy, uz;
};

// 1/x: one v_rcp_f32 (1 ulp) in float, exact division in double
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// Sums start from -0, the additive identity of IEEE arithmetic (x + -0 = x for
// every x, +0 included; +0 is not one, since -0 + +0 = +0), so the first add or
// fma of a sum folds into a move or a multiply.
template <typename C>
constexpr C kSum0 = C(-0.0);

// a sign-folded term: acc + c v for c in {-1, 1}
template <typename C>
__device__ __forceinline__ C add_signed(C acc, int c, C v) {
  return c > 0 ? acc + v : acc - v;
}

template <class L>
constexpr bool has_z() {
  for (int q = 0; q < L::Q; ++q)
    if (L::cz[q] != 0) return true;
  return false;
}

// rho = sum f, u = sum c f / max(rho, 1e-12)   (d3q19.py:165-186).  Summed one
// population at a time in q order: every kernel (one step, two steps per launch,
// the closure march) then rounds these sums identically.  A sum over opposite
// pairs (f_q + f_opp, f_q - f_opp) saves a few adds but made the one- and
// two-step kernels disagree by an ulp in about 1 cell in 5000 of masked grids.
// kDevIn: f holds deviations f - w_q, whose sum is rho - 1 (the weights sum to 1
// and their first moments vanish, so the momentum sums are unchanged).
template <class L, typename C, bool kDevIn = false>
__device__ __forceinline__ Macro<C> macro(const C (&f)[L::Q]) {
  C rho = kSum0<C>, jx = kSum0<C>, jy = kSum0<C>, jz = kSum0<C>;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    rho += f[q];
    if (L::cx[q] != 0) jx = add_signed(jx, L::cx[q], f[q]);
    if (L::cy[q] != 0) jy = add_signed(jy, L::cy[q], f[q]);
    if (L::cz[q] != 0) jz = add_signed(jz, L::cz[q], f[q]);
  }
  if constexpr (!has_z<L>()) jz = C(0);
  if constexpr (kDevIn) rho += C(1);
  Macro<C> m;
  m.rho = rho;
  m.rho_safe = rho > C(1e-12) ? rho : C(1e-12);
  const C inv = recip(m.rho_safe);
  m.ux = jx * inv;
  m.uy = jy * inv;
  m.uz = jz * inv;
  return m;
}

// w rho (1 + 3 cu + 4.5 cu^2 - 1.5 u^2)   (d3q19.py:119-121)
template <class L, typename C>
__device__ __forceinline__ void equilibrium(C rho, C ux, C uy, C uz, C (&feq)[L::Q]) {
  const C usq15 = C(1.5) * (ux * ux + uy * uy + uz * uz);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    C cu = kSum0<C>;
    if (L::cx[q] != 0) cu += C(L::cx[q]) * ux;
    if (L::cy[q] != 0) cu += C(L::cy[q]) * uy;
    if (L::cz[q] != 0) cu += C(L::cz[q]) * uz;
    feq[q] = C(L::w[q]) * rho * (C(1) + C(3) * cu + C(4.5) * cu * cu - usq15);
  }
}

// BGK relaxation f <- (1-w) f + w f_eq with the equilibrium evaluated per
// opposite pair: f_eq(q), f_eq(opp q) = s +/- a with s = w_q rho (1 - 1.5u^2 + 4.5 cu^2)
// and a = 3 w_q rho cu (same algebra as d3q19.py:119-121, fewer operations).
// kDevOut returns f - w_q instead, the deviation storage's value: s - w_q =
// w_q (base - 1) + ..., one subtraction per cell in place of one per population.
// kDevIn (with kDevOut) takes f - w_q as well: (1 - omega) w_q joins the same
// term, w_q (base - omega).
template <class L, typename C, bool kDevOut = false, bool kDevIn = false>
__device__ __forceinline__ void bgk_relax(C (&f)[L::Q], C rho, C ux, C uy, C uz, C omega) {
  static_assert(kDevOut || !kDevIn, "deviations in, deviations out");
  const C om1 = C(1) - omega;
  const C orho = omega * rho;
  const C base0 = orho * (C(1) - C(1.5) * (ux * ux + uy * uy + uz * uz));
  const C base = kDevIn ? base0 - omega : (kDevOut ? base0 - C(1) : base0);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int o = L::opp[q];
    if (o == q) {
      f[q] = om1 * f[q] + C(L::w[q]) * base;
    } else if (q < o) {
      C cu = kSum0<C>;
      if (L::cx[q] != 0) cu += C(L::cx[q]) * ux;
      if (L::cy[q] != 0) cu += C(L::cy[q]) * uy;
      if (L::cz[q] != 0) cu += C(L::cz[q]) * uz;
      const C k = C(L::w[q]) * orho;
      const C sym = (C(4.5) * k) * cu * cu + C(L::w[q]) * base;
      const C asym = (C(3) * k) * cu;
      f[q] = om1 * f[q] + (sym + asym);
      f[o] = om1 * f[o] + (sym - asym);
    }
  }
}

// Zou-He moving lid on the y = ny-1 plane, ux = u, uy = uz = 0 (boundaries3d.py:324-404)
template <typename C>
__device__ __forceinline__ void zou_he_lid(C (&f)[19], C u) {
  const C s0 = f[0] + f[1] + f[2] + f[5] + f[6] + f[11] + f[12] + f[13] + f[14];
  const C sp = f[3] + f[7] + f[10] + f[15] + f[17];
  const C rho = s0 + C(2) * sp;
  const C dx = rho * u - ((f[1] - f[2] + f[11] - f[12] + f[13] - f[14]) + (f[7] - f[10]));
  const C sz = f[5] - f[6] + f[11] - f[12] - f[13] + f[14];
  const C h = C(0.5) * (f[7] + f[10]);
  f[4] = f[3];
  f[8] = h - C(0.5) * dx;
  f[9] = h + C(0.5) * dx;
  f[16] = f[15] + C(0.5) * sz;
  f[18] = f[17] - C(0.5) * sz;
}

// Second-order non-equilibrium stress Pi_ab = sum c_a c_b fneq (xx yy zz xy xz yz)
template <class L, typename C>
__device__ __forceinline__ void neq_stress(const C (&fneq)[L::Q], C (&pi)[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) pi[k] = kSum0<C>;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    if (L::cx[q] != 0) pi[0] += fneq[q];
    if (L::cy[q] != 0) pi[1] += fneq[q];
    if (L::cz[q] != 0) pi[2] += fneq[q];
    if (L::cx[q] * L::cy[q] != 0) pi[3] += C(L::cx[q] * L::cy[q]) * fneq[q];
    if (L::cx[q] * L::cz[q] != 0) pi[4] += C(L::cx[q] * L::cz[q]) * fneq[q];
    if (L::cy[q] * L::cz[q] != 0) pi[5] += C(L::cy[q] * L::cz[q]) * fneq[q];
  }
}

// Pi^neq = sum_q c_a c_b f_q - rho (delta_ab / 3 + u_a u_b): the second moments of
// the second-order equilibrium are exact on lattices with 4th-order isotropic
// weights (D2Q9, D3Q19, D3Q27), so this equals neq_stress(f - f_eq) without
// forming f_eq.
// kDevIn: f holds deviations f - w_q, whose diagonal second moments lack the
// weights' 1/3, so rho / 3 becomes (rho - 1) / 3.
template <class L, typename C, bool kDevIn = false>
__device__ __forceinline__ void neq_stress_moments(const C (&f)[L::Q], C rho, C ux, C uy, C uz, C (&pi)[6]) {
  // second moments from the opposite-pair sums f_q + f_opp (c_a c_b is even in c)
  C m[6] = {kSum0<C>, kSum0<C>, kSum0<C>, kSum0<C>, kSum0<C>, kSum0<C>};
  {
#pragma clang fp contract(off)  // no fusion into the pair sums, whatever the calling kernel
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      const int o = L::opp[q];
      if (q >= o) continue;
      const C sq = f[q] + f[o];
      if (L::cx[q] != 0) m[0] += sq;
      if (L::cy[q] != 0) m[1] += sq;
      if (L::cz[q] != 0) m[2] += sq;
      if (L::cx[q] * L::cy[q] != 0) m[3] = add_signed(m[3], L::cx[q] * L::cy[q], sq);
      if (L::cx[q] * L::cz[q] != 0) m[4] = add_signed(m[4], L::cx[q] * L::cz[q], sq);
      if (L::cy[q] * L::cz[q] != 0) m[5] = add_signed(m[5], L::cy[q] * L::cz[q], sq);
    }
  }
  if constexpr (!has_z<L>()) m[2] = m[4] = m[5] = C(0);
  const C r3 = (kDevIn ? rho - C(1) : rho) * C(1.0 / 3.0);
  pi[0] = m[0] - (r3 + rho * ux * ux);
  pi[1] = m[1] - (r3 + rho * uy * uy);
  pi[2] = m[2] - (r3 + rho * uz * uz);
  pi[3] = m[3] - rho * ux * uy;
  pi[4] = m[4] - rho * ux * uz;
  pi[5] = m[5] - rho * uy * uz;
}

template <typename C>
__device__ __forceinline__ C pi_norm(const C (&pi)[6]) {
  return sqrt(pi[0] * pi[0] + pi[1] * pi[1] + pi[2] * pi[2] +
              C(2) * (pi[3] * pi[3] + pi[4] * pi[4] + pi[5] * pi[5]));
}

// Hou et al. tau_eff clamped to [0.5001, max(1, tau+0.5)]  (turbulence.py:183-208).
// tau^2 enters through an explicit fma: as a product of launch constants it would
// otherwise be fused or not depending on whether the compiler hoists it out of a
// loop, and kernels sharing this code must round identically.
template <typename C>
__device__ __forceinline__ C smagorinsky_tau(C pn, C rho_safe, C tau, C cs_delta_sq, C tau_cap) {
  C disc = fma(tau, tau, C(18) * cs_delta_sq * pn / rho_safe);
  disc = disc > C(0) ? disc : C(0);
  C te = C(0.5) * (tau + sqrt(disc));
  te = te < C(0.5001) ? C(0.5001) : te;
  return te > tau_cap ? tau_cap : te;
}

template <typename C>
__device__ __forceinline__ C smagorinsky_omega(C pn, C rho_safe, C tau, C cs_delta_sq, C tau_cap) {
  return C(1) / smagorinsky_tau(pn, rho_safe, tau, cs_delta_sq, tau_cap);
}

// Unclamped closure of collide_cumulant_d3q19 (cumulant.py:1316-1321)
template <typename C>
__device__ __forceinline__ C smagorinsky_raw_omega(C pn, C rho_safe, C tau, C cs_delta_sq) {
  return C(2) / (tau + sqrt(fma(tau, tau, C(18) * cs_delta_sq * pn / rho_safe)));
}

// 4.5 w_q (H_q : Pi) with doubled off-diagonals (cumulant.py:1331-1362)
template <class L, typename C>
__device__ __forceinline__ C hermite2(int q, const C (&pi)[6]) {
  const C t = C(1.0 / 3.0);
  const C h = (C(L::cx[q] * L::cx[q]) - t) * pi[0] + (C(L::cy[q] * L::cy[q]) - t) * pi[1] +
              (C(L::cz[q] * L::cz[q]) - t) * pi[2] + C(2 * L::cx[q] * L::cy[q]) * pi[3] +
              C(2 * L::cx[q] * L::cz[q]) * pi[4] + C(2 * L::cy[q] * L::cz[q]) * pi[5];
  return C(4.5) * C(L::w[q]) * h;
}

template <class L>
struct MrtTab;
template <>
struct MrtTab<D2Q9> {
  static constexpr const double (&M)[9][9] = tables::MRT9;
  static constexpr const double (&Mi)[9][9] = tables::MRT9_INV;
};
// D3Q19 / D3Q27 also carry the MRT basis as a combination of the monomial
// moments: M = T CM and M^-1 = CM^-1 TI (CM: the monomial moment matrix).
template <>
struct MrtTab<D3Q19> {
  static constexpr const double (&M)[19][19] = tables::MRT19;
  static constexpr const double (&Mi)[19][19] = tables::MRT19_INV;
  static constexpr const double (&T)[19][19] = tables::MRT19_T;
  static constexpr const double (&TI)[19][19] = tables::MRT19_TI;
};
template <>
struct MrtTab<D3Q27> {
  static constexpr const double (&M)[27][27] = tables::MRT27;
  static constexpr const double (&Mi)[27][27] = tables::MRT27_INV;
  static constexpr const double (&T)[27][27] = tables::MRT27_T;
  static constexpr const double (&TI)[27][27] = tables::MRT27_TI;
};

// monomial moments and their inverse (defined with the central-moment tables below)
template <class L, typename C>
__device__ __forceinline__ void moments(const C (&v)[L::Q], C (&m)[L::Q]);
template <class L, typename C>
__device__ __forceinline__ void populations(const C (&m)[L::Q], C (&v)[L::Q]);

// f* = f - M^-1 S M (f - f_eq); S holds the per-moment rates, with the shear
// entries given by omega (may vary per cell).  solver3d.py:107-160.
// Relaxation group of MRT moment i (rows of MrtTab<L>::M; rates in the order the
// reference's s vectors list them, solver.py:249-335, solver3d.py:64-90, d3q27.py:643-726):
// conserved moments are never relaxed, the shear rows take omega.
enum MrtGroup { MRT_CONSERVED, MRT_SE, MRT_SEPS, MRT_SQ, MRT_SPI, MRT_ONE, MRT_SHEAR };

template <class L>
constexpr int mrt_group(int i) {
  if (L::Q == 9) {
    constexpr int g[9] = {MRT_CONSERVED, MRT_SE, MRT_SEPS, MRT_CONSERVED, MRT_SQ, MRT_CONSERVED, MRT_SQ,
                          MRT_SHEAR, MRT_SHEAR};
    return g[i];
  } else if (L::Q == 19) {
    constexpr int g[19] = {MRT_CONSERVED, MRT_SE, MRT_SEPS, MRT_CONSERVED, MRT_SQ, MRT_CONSERVED, MRT_SQ,
                           MRT_CONSERVED, MRT_SQ, MRT_SHEAR, MRT_SHEAR, MRT_SHEAR, MRT_SHEAR, MRT_SHEAR,
                           MRT_SPI, MRT_SPI, MRT_ONE, MRT_ONE, MRT_ONE};
    return g[i];
  } else {
    return i < 4 ? MRT_CONSERVED
                 : (i == 4 ? MRT_SE : (i < 10 ? MRT_SHEAR : (i < 19 ? MRT_SQ : (i == 19 ? MRT_SEPS : MRT_SPI))));
  }
}

// rate = (s_e, s_eps, s_q, s_pi) of the collision parameters
template <class L, typename C>
__device__ __forceinline__ C mrt_rate(int i, const C* rate, C omega) {
  const int g = mrt_group<L>(i);
  return g == MRT_SE     ? rate[0]
         : g == MRT_SEPS ? rate[1]
         : g == MRT_SQ   ? rate[2]
         : g == MRT_SPI  ? rate[3]
         : g == MRT_ONE  ? C(1)
                         : omega;
}

template <class L, typename C>
__device__ __forceinline__ void mrt(C (&f)[L::Q], const C (&feq)[L::Q], const C* rate, C omega) {
  constexpr int Q = L::Q;
  C m[Q];
  if constexpr (Q == 19 || Q == 27) {
    // through the monomial moments: axis-factored transforms and the sparse T / TI
    C v[Q], mono[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = f[q] - feq[q];
    moments<L>(v, mono);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      m[i] = C(0);
      if (mrt_group<L>(i) == MRT_CONSERVED) continue;  // relaxed at rate 0: never formed
      C acc = kSum0<C>;
#pragma unroll
      for (int k = 0; k < Q; ++k)
        if (MrtTab<L>::T[i][k] != 0.0) acc += C(MrtTab<L>::T[i][k]) * mono[k];
      m[i] = acc * mrt_rate<L>(i, rate, omega);
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      C acc = kSum0<C>;
#pragma unroll
      for (int i = 0; i < Q; ++i)
        if (mrt_group<L>(i) != MRT_CONSERVED && MrtTab<L>::TI[k][i] != 0.0) acc += C(MrtTab<L>::TI[k][i]) * m[i];
      mono[k] = acc;
    }
    populations<L>(mono, v);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] -= v[q];
    return;
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    m[i] = C(0);
    if (mrt_group<L>(i) == MRT_CONSERVED) continue;  // relaxed at rate 0: never formed
    C acc = kSum0<C>;
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      constexpr double zero = 0.0;
      if (MrtTab<L>::M[i][j] != zero) acc += C(MrtTab<L>::M[i][j]) * (f[j] - feq[j]);
    }
    const int g = mrt_group<L>(i);
    const C r = g == MRT_SE     ? rate[0]
                : g == MRT_SEPS ? rate[1]
                : g == MRT_SQ   ? rate[2]
                : g == MRT_SPI  ? rate[3]
                : g == MRT_ONE  ? C(1)
                                : omega;
    m[i] = acc * r;
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    C acc = kSum0<C>;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      if (mrt_group<L>(i) != MRT_CONSERVED && MrtTab<L>::Mi[q][i] != 0.0) acc += C(MrtTab<L>::Mi[q][i]) * m[i];
    }
    f[q] -= acc;
  }
}

template <class L>
struct CmTab;
template <>
struct CmTab<D2Q9> {
  static constexpr const double (&M)[9][9] = tables::CM9;
  static constexpr const double (&Mi)[9][9] = tables::CM9_INV;
  static constexpr const int (&SX)[3][3] = tables::CM9_SHIFT_X;
  static constexpr const int (&SY)[3][3] = tables::CM9_SHIFT_Y;
  static constexpr int NG = 3;
};
template <>
struct CmTab<D3Q19> {
  static constexpr const double (&M)[19][19] = tables::CM19;
  static constexpr const double (&Mi)[19][19] = tables::CM19_INV;
  static constexpr const int (&SX)[5][3] = tables::CM19_SHIFT_X;
  static constexpr const int (&SY)[5][3] = tables::CM19_SHIFT_Y;
  static constexpr const int (&SZ)[5][3] = tables::CM19_SHIFT_Z;
  static constexpr int NG = 5;
};
template <>
struct CmTab<D3Q27> {
  static constexpr const double (&M)[27][27] = tables::CM27;
  static constexpr const double (&Mi)[27][27] = tables::CM27_INV;
  static constexpr const int (&SX)[9][3] = tables::CM27_SHIFT_X;
  static constexpr const int (&SY)[9][3] = tables::CM27_SHIFT_Y;
  static constexpr const int (&SZ)[9][3] = tables::CM27_SHIFT_Z;
  static constexpr int NG = 9;
};

// binomial 1-D shift, sign -1 forward (raw -> central), +1 inverse
template <int NG, typename C, int Q>
__device__ __forceinline__ void shift1d(C (&m)[Q], const int (&g)[NG][3], C u, C sign) {
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    const C m0 = m[g[k][0]], m1 = m[g[k][1]], m2 = m[g[k][2]];
    m[g[k][1]] = m1 + sign * u * m0;
    m[g[k][2]] = m2 + sign * C(2) * u * m1 + u * u * m0;
  }
}

// D3Q27 is the full 3^3 velocity set, so its monomial moments factor by axis:
// along each axis a line of three values (c = -1, 0, +1) maps to its moments of
// degree 0, 1, 2 as  s = v+ + v-,  m0 = s + v0,  m1 = v+ - v-,  m2 = s,  and
// back as  v0 = m0 - m2,  v+ = (m2 + m1) / 2,  v- = (m2 - m1) / 2.  Three sweeps
// of nine lines each: 81 operations forward and 108 back, against ~343 for the
// 27 x 27 table products.  Moment order: the exponents (a, b, c) of the CM27
// rows (gen_tables.py DEG27); population order: D3Q27's velocities.
namespace f27 {
constexpr int DEG[27][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
                            {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1},
                            {1, 0, 2}, {0, 1, 2}, {1, 1, 1}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1},
                            {1, 2, 1}, {1, 1, 2}, {2, 2, 1}, {2, 1, 2}, {1, 2, 2}, {2, 2, 2}};
struct Index {
  int pop[3][3][3];  // [cx + 1][cy + 1][cz + 1] -> population
  int mom[3][3][3];  // [a][b][c] -> moment
};
constexpr Index make_index() {
  Index t{};
  for (int q = 0; q < 27; ++q) t.pop[D3Q27::cx[q] + 1][D3Q27::cy[q] + 1][D3Q27::cz[q] + 1] = q;
  for (int i = 0; i < 27; ++i) t.mom[DEG[i][0]][DEG[i][1]][DEG[i][2]] = i;
  return t;
}
constexpr Index IDX = make_index();
constexpr double ipow(int c, int e) { return e == 0 ? 1.0 : (e == 1 ? double(c) : double(c * c)); }
// the exponent table is the one the CM27 rows were built from
constexpr bool matches_table() {
  for (int i = 0; i < 27; ++i)
    for (int q = 0; q < 27; ++q)
      if (tables::CM27[i][q] != ipow(D3Q27::cx[q], DEG[i][0]) * ipow(D3Q27::cy[q], DEG[i][1]) *
                                   ipow(D3Q27::cz[q], DEG[i][2]))
        return false;
  return true;
}
static_assert(matches_table(), "D3Q27 moment exponents must follow the CM27 rows");

// Sweep steps that know at compile time which operands are structurally zero
// (the D3Q19 populations and moments embedded in D3Q27 arrays leave 8 of 27
// entries zero): an operand flagged zero is dropped instead of added, since
// x + 0 does not fold for the compiler (it differs from x when x = -0).  The
// flags are constants once the sweeps are unrolled; the results equal the
// plain sums except possibly for the sign of an exact zero.
template <typename C>
__device__ __forceinline__ C zadd(C a, bool za, C b, bool zb) {
  return za ? b : (zb ? a : a + b);
}
template <typename C>
__device__ __forceinline__ C zsub(C a, bool za, C b, bool zb) {
  return zb ? a : (za ? -b : a - b);
}
template <typename C>
__device__ __forceinline__ C zfma(C k, C a, bool za, C b, bool zb) {
  return za ? b : (zb ? k * a : fma(k, a, b));
}

// raw moments m[DEG] = sum_q cx^a cy^b cz^c v_q by three 1-D sweeps (x, y, z);
// zv[q] flags the populations known to be zero
template <typename C>
__device__ __forceinline__ void forward(const C (&v)[27], C (&m)[27], const bool (&zv)[27]) {
  C g[3][3][3], h[3][3][3];  // after the x sweep: [a][cy][cz]; after the y sweep: [a][b][cz]
  bool zg[3][3][3], zh[3][3][3];
#pragma unroll
  for (int y = 0; y < 3; ++y)
#pragma unroll
    for (int z = 0; z < 3; ++z) {
      const int qm = IDX.pop[0][y][z], q0 = IDX.pop[1][y][z], qp = IDX.pop[2][y][z];
      const bool zs = zv[qp] && zv[qm];
      const C sum = zadd(v[qp], zv[qp], v[qm], zv[qm]);
      g[0][y][z] = zadd(sum, zs, v[q0], zv[q0]);
      g[1][y][z] = zsub(v[qp], zv[qp], v[qm], zv[qm]);
      g[2][y][z] = sum;
      zg[0][y][z] = zs && zv[q0];
      zg[1][y][z] = zg[2][y][z] = zs;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int z = 0; z < 3; ++z) {
      const bool zm = zg[a][0][z], z0 = zg[a][1][z], zp = zg[a][2][z], zs = zp && zm;
      const C sum = zadd(g[a][2][z], zp, g[a][0][z], zm);
      h[a][0][z] = zadd(sum, zs, g[a][1][z], z0);
      h[a][1][z] = zsub(g[a][2][z], zp, g[a][0][z], zm);
      h[a][2][z] = sum;
      zh[a][0][z] = zs && z0;
      zh[a][1][z] = zh[a][2][z] = zs;
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const bool zm = zh[a][b][0], z0 = zh[a][b][1], zp = zh[a][b][2], zs = zp && zm;
      const C sum = zadd(h[a][b][2], zp, h[a][b][0], zm);
      m[IDX.mom[a][b][0]] = zadd(sum, zs, h[a][b][1], z0);
      m[IDX.mom[a][b][1]] = zsub(h[a][b][2], zp, h[a][b][0], zm);
      m[IDX.mom[a][b][2]] = sum;
    }
}

// the populations of raw moments, the three sweeps inverted; zm[i] flags the
// moments known to be zero
template <typename C>
__device__ __forceinline__ void inverse(const C (&m)[27], C (&v)[27], const bool (&zm)[27]) {
  C g[3][3][3], h[3][3][3];  // [a][b][cz], then [a][cy][cz]
  bool zg[3][3][3], zh[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int i0 = IDX.mom[a][b][0], i1 = IDX.mom[a][b][1], i2 = IDX.mom[a][b][2];
      const C half = C(0.5) * m[i2];
      h[a][b][1] = zsub(m[i0], zm[i0], m[i2], zm[i2]);
      h[a][b][2] = zfma(C(0.5), m[i1], zm[i1], half, zm[i2]);
      h[a][b][0] = zfma(C(-0.5), m[i1], zm[i1], half, zm[i2]);
      zh[a][b][1] = zm[i0] && zm[i2];
      zh[a][b][0] = zh[a][b][2] = zm[i1] && zm[i2];
    }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int z = 0; z < 3; ++z) {
      const bool z0 = zh[a][0][z], z1 = zh[a][1][z], z2 = zh[a][2][z];
      const C half = C(0.5) * h[a][2][z];
      g[a][1][z] = zsub(h[a][0][z], z0, h[a][2][z], z2);
      g[a][2][z] = zfma(C(0.5), h[a][1][z], z1, half, z2);
      g[a][0][z] = zfma(C(-0.5), h[a][1][z], z1, half, z2);
      zg[a][1][z] = z0 && z2;
      zg[a][0][z] = zg[a][2][z] = z1 && z2;
    }
#pragma unroll
  for (int y = 0; y < 3; ++y)
#pragma unroll
    for (int z = 0; z < 3; ++z) {
      const bool z0 = zg[0][y][z], z1 = zg[1][y][z], z2 = zg[2][y][z];
      const C half = C(0.5) * g[2][y][z];
      v[IDX.pop[1][y][z]] = zsub(g[0][y][z], z0, g[2][y][z], z2);
      v[IDX.pop[2][y][z]] = zfma(C(0.5), g[1][y][z], z1, half, z2);
      v[IDX.pop[0][y][z]] = zfma(C(-0.5), g[1][y][z], z1, half, z2);
    }
}

constexpr bool kNone27[27] = {};

template <typename C>
__device__ __forceinline__ void forward(const C (&v)[27], C (&m)[27]) {
  forward(v, m, kNone27);
}

template <typename C>
__device__ __forceinline__ void inverse(const C (&m)[27], C (&v)[27]) {
  inverse(m, v, kNone27);
}
// Raw moments of the second-order equilibrium w_q rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2)
// in closed form.  D3Q27's weights are products of 1-D weights (2/3, 1/6, 1/6) whose
// moments of degree 0..4 are 1, 0, 1/3, 0, 1/3, so with S(e) = M(e_x) M(e_y) M(e_z):
//   m_e = rho [ (1 - 1.5 u^2) S(e) + 3 sum_d u_d S(e + d) + 4.5 sum_{d,f} u_d u_f S(e + d + f) ].
constexpr double m1d(int n) { return n == 0 ? 1.0 : (n % 2 ? 0.0 : 1.0 / 3.0); }
constexpr double s3(int a, int b, int c) { return m1d(a) * m1d(b) * m1d(c); }
struct EqCoef {
  double c0[27];     // of (1 - 1.5 u^2)
  double c1[27][3];  // of u_d
  double c2[27][6];  // of u_x^2, u_y^2, u_z^2, u_x u_y, u_x u_z, u_y u_z
};
constexpr EqCoef make_eq_coef() {
  EqCoef t{};
  constexpr int pd[6] = {0, 1, 2, 0, 0, 1}, pf[6] = {0, 1, 2, 1, 2, 2};
  for (int i = 0; i < 27; ++i) {
    const int a = DEG[i][0], b = DEG[i][1], c = DEG[i][2];
    t.c0[i] = s3(a, b, c);
    for (int d = 0; d < 3; ++d) t.c1[i][d] = 3.0 * s3(a + (d == 0), b + (d == 1), c + (d == 2));
    for (int k = 0; k < 6; ++k) {
      const int d = pd[k], f = pf[k];
      t.c2[i][k] = (d == f ? 4.5 : 9.0) *
                   s3(a + (d == 0) + (f == 0), b + (d == 1) + (f == 1), c + (d == 2) + (f == 2));
    }
  }
  return t;
}
constexpr EqCoef EQ = make_eq_coef();

template <typename C>
__device__ __forceinline__ void eq_moments(C rho, C ux, C uy, C uz, C (&m)[27]) {
  const C u[3] = {ux, uy, uz};
  const C uu[6] = {ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz};
  const C base = C(1) - C(1.5) * (uu[0] + uu[1] + uu[2]);
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    C acc = EQ.c0[i] != 0.0 ? C(EQ.c0[i]) * base : C(0);
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (EQ.c1[i][d] != 0.0) acc = fma(C(EQ.c1[i][d]), u[d], acc);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (EQ.c2[i][k] != 0.0) acc = fma(C(EQ.c2[i][k]), uu[k], acc);
    m[i] = rho * acc;
  }
}
// D3Q19 through the same sweeps: its velocities are D3Q27's without the eight
// corners and its moments a subset of D3Q27's exponents.  A distribution without
// corner populations has zero moments wherever all three exponents are nonzero
// (only corners have three nonzero components), so its D3Q27 moments are its
// D3Q19 moments plus those zeros, and the D3Q27 inverse of that vector is the
// distribution itself, corners zero.  Exponent order: the CM19 rows.
constexpr int DEG19[19][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
                              {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1},
                              {1, 0, 2}, {0, 1, 2}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}};
constexpr bool matches_table19() {
  for (int i = 0; i < 19; ++i)
    for (int q = 0; q < 19; ++q)
      if (tables::CM19[i][q] != ipow(D3Q19::cx[q], DEG19[i][0]) * ipow(D3Q19::cy[q], DEG19[i][1]) *
                                   ipow(D3Q19::cz[q], DEG19[i][2]))
        return false;
  return true;
}
static_assert(matches_table19(), "D3Q19 moment exponents must follow the CM19 rows");

// the D3Q27 entries a D3Q19 distribution leaves zero: the 8 corner populations
// and the 8 moments whose three exponents are all nonzero
struct Zero19 {
  bool pop[27];
  bool mom[27];
};
constexpr Zero19 make_zero19() {
  Zero19 t{};
  for (int q = 0; q < 27; ++q) t.pop[q] = D3Q27::cx[q] != 0 && D3Q27::cy[q] != 0 && D3Q27::cz[q] != 0;
  for (int i = 0; i < 27; ++i) t.mom[i] = DEG[i][0] != 0 && DEG[i][1] != 0 && DEG[i][2] != 0;
  return t;
}
constexpr Zero19 Z19 = make_zero19();

template <typename C>
__device__ __forceinline__ void forward19(const C (&v)[19], C (&m)[19]) {
  C v27[27], m27[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) v27[q] = C(0);
#pragma unroll
  for (int q = 0; q < 19; ++q) v27[IDX.pop[D3Q19::cx[q] + 1][D3Q19::cy[q] + 1][D3Q19::cz[q] + 1]] = v[q];
  forward(v27, m27, Z19.pop);
#pragma unroll
  for (int i = 0; i < 19; ++i) m[i] = m27[IDX.mom[DEG19[i][0]][DEG19[i][1]][DEG19[i][2]]];
}

template <typename C>
__device__ __forceinline__ void inverse19(const C (&m)[19], C (&v)[19]) {
  C m27[27], v27[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) m27[i] = C(0);
#pragma unroll
  for (int i = 0; i < 19; ++i) m27[IDX.mom[DEG19[i][0]][DEG19[i][1]][DEG19[i][2]]] = m[i];
  inverse(m27, v27, Z19.mom);
#pragma unroll
  for (int q = 0; q < 19; ++q) v[q] = v27[IDX.pop[D3Q19::cx[q] + 1][D3Q19::cy[q] + 1][D3Q19::cz[q] + 1]];
}
}  // namespace f27

template <class L, typename C>
__device__ __forceinline__ void moments(const C (&v)[L::Q], C (&m)[L::Q]) {
  if constexpr (L::Q == 27) {
    f27::forward(v, m);
    return;
  }
  if constexpr (L::Q == 19) {
    f27::forward19(v, m);
    return;
  }
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    C acc = kSum0<C>;
#pragma unroll
    for (int j = 0; j < L::Q; ++j)
      if (CmTab<L>::M[i][j] != 0.0) acc += C(CmTab<L>::M[i][j]) * v[j];
    m[i] = acc;
  }
}

template <class L, typename C>
__device__ __forceinline__ void populations(const C (&m)[L::Q], C (&v)[L::Q]) {
  if constexpr (L::Q == 27) {
    f27::inverse(m, v);
    return;
  }
  if constexpr (L::Q == 19) {
    f27::inverse19(m, v);
    return;
  }
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    C acc = kSum0<C>;
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
      if (CmTab<L>::Mi[q][i] != 0.0) acc += C(CmTab<L>::Mi[q][i]) * m[i];
    v[q] = acc;
  }
}

template <class L, typename C>
__device__ __forceinline__ void to_central(C (&m)[L::Q], C ux, C uy, C uz) {
  shift1d<CmTab<L>::NG>(m, CmTab<L>::SX, ux, C(-1));
  shift1d<CmTab<L>::NG>(m, CmTab<L>::SY, uy, C(-1));
  if constexpr (L::Q != 9) shift1d<CmTab<L>::NG>(m, CmTab<L>::SZ, uz, C(-1));
}

template <class L, typename C>
__device__ __forceinline__ void to_raw(C (&m)[L::Q], C ux, C uy, C uz) {
  if constexpr (L::Q != 9) shift1d<CmTab<L>::NG>(m, CmTab<L>::SZ, uz, C(1));
  shift1d<CmTab<L>::NG>(m, CmTab<L>::SY, uy, C(1));
  shift1d<CmTab<L>::NG>(m, CmTab<L>::SX, ux, C(1));
}

// Cascaded central-moment collision on f_neq (cascaded_collision.py:445-509).
// rate = {s_bulk, s_3, s_4, s_5, s_6}.  On D2Q9 this is collide_cumulant_d2q9
// (cumulant.py:116-311: trace at omega_b, deviatoric stress at omega, kappa21 /
// kappa12 at omega_3, kappa22 at omega_4), rate = {omega_b, omega_3, omega_4}.
// Relaxing the central moments of f_neq = f - f_eq at rate s is relaxing those
// of f towards the central moments of f_eq, since the shift is linear in f.
template <class L, typename C>
__device__ __forceinline__ void cascaded(C (&f)[L::Q], const C (&feq)[L::Q], const Macro<C>& mc,
                                         C omega, const C* rate) {
  constexpr int Q = L::Q;
  C v[Q], m[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = f[q] - feq[q];
  moments<L>(v, m);
  to_central<L>(m, mc.ux, mc.uy, mc.uz);
  if constexpr (Q == 9) {
    // m: 00 10 01 20 02 11 21 12 22
    const C tr = m[3] + m[4];
    const C trs = (C(1) - rate[0]) * tr;
    m[3] = (C(1) - omega) * (m[3] - C(0.5) * tr) + C(0.5) * trs;
    m[4] = (C(1) - omega) * (m[4] - C(0.5) * tr) + C(0.5) * trs;
    m[5] = (C(1) - omega) * m[5];
    m[6] = (C(1) - rate[1]) * m[6];
    m[7] = (C(1) - rate[1]) * m[7];
    m[8] = (C(1) - rate[2]) * m[8];
    to_raw<L>(m, mc.ux, mc.uy, mc.uz);
    populations<L>(m, v);
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = feq[q] + v[q];
    return;
  }
  const C tr = m[4] + m[5] + m[6];
  const C trs = (C(1) - rate[0]) * tr;
  const C third = C(1.0 / 3.0);
#pragma unroll
  for (int i = 4; i < 7; ++i) m[i] = (C(1) - omega) * (m[i] - tr * third) + trs * third;
#pragma unroll
  for (int i = 7; i < 10; ++i) m[i] = (C(1) - omega) * m[i];
  if (Q == 19) {
#pragma unroll
    for (int i = 10; i < 16; ++i) m[i] = (C(1) - rate[1]) * m[i];
#pragma unroll
    for (int i = 16; i < Q; ++i) m[i] = (C(1) - rate[2]) * m[i];
  } else {
#pragma unroll
    for (int i = 10; i < 17; ++i) m[i] = (C(1) - rate[1]) * m[i];
#pragma unroll
    for (int i = 17; i < 23; ++i) m[i] = (C(1) - rate[2]) * m[i];
#pragma unroll
    for (int i = 23; i < 26; ++i) m[i] = (C(1) - rate[3]) * m[i];
#pragma unroll
    for (int i = 26; i < Q; ++i) m[i] = (C(1) - rate[4]) * m[i];
  }
  to_raw<L>(m, mc.ux, mc.uy, mc.uz);
  populations<L>(m, v);
#pragma unroll
  for (int q = 0; q < Q; ++q) f[q] = feq[q] + v[q];
}

// D3Q19 Hermite-regularised cumulant (cumulant.py:1159-1377).
// rate = {omega_b, omega_odd (unused), omega_even}.
template <class L, typename C>
__device__ __forceinline__ void cumulant19(C (&f)[L::Q], const C (&feq)[L::Q], const C (&pi)[6],
                                           C omega, const C* rate) {
  const C ob = rate[0], oe = rate[2];
  const C tr = pi[0] + pi[1] + pi[2];
  const C db = (ob - omega) * tr / C(3);
  C ps[6];
  ps[0] = pi[0] - omega * pi[0] - db;
  ps[1] = pi[1] - omega * pi[1] - db;
  ps[2] = pi[2] - omega * pi[2] - db;
  ps[3] = pi[3] - omega * pi[3];
  ps[4] = pi[4] - omega * pi[4];
  ps[5] = pi[5] - omega * pi[5];
  // f_eq + H(ps) + (1 - oe) (f - f_eq - H(pi)) with H linear in its tensor:
  // f + oe (f_eq - f) + H(ps - (1 - oe) pi), one Hermite contraction per population
  const C keep = C(1) - oe;
  C d[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) d[k] = ps[k] - keep * pi[k];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) f[q] = fma(oe, feq[q] - f[q], f[q]) + hermite2<L>(q, d);
}

// Geier 2015 cumulants on the full distribution, relaxed towards the
// equilibrium's cumulants (cumulant.py:1050-1156).  rate = {omega_b,
// omega_odd, omega_even}.
// Cumulants C = rho C~(kappa / rho) of the second-order equilibrium at density rho,
// in the DEG27 order, in closed form: its raw moments (f27::eq_moments), shifted
// to central moments and mapped by the set-partition formula, reduce to
//   C_200 = rho/3,  C_110 = 0,  C_210 = -rho ux^2 uy,  C_111 = -rho ux uy uz,
//   C_220 = 3 rho ux^2 uy^2,  C_211 = 3 rho ux^2 uy uz,  C_221 = -6 rho ux^2 uy^2 uz,
//   C_222 = 0  (and permutations), i.e. the product of that pipeline, which they
// replace (checked against it in floating point and symbolically in the tests).
template <typename C>
__device__ __forceinline__ void eq_cumulants27(C rho, C ux, C uy, C uz, C (&ce)[27]) {
  const C rx2 = rho * ux * ux, ry2 = rho * uy * uy, rz2 = rho * uz * uz, rxyz = rho * ux * uy * uz;
  const C r3 = rho * C(1.0 / 3.0);
  ce[0] = rho;
  ce[1] = ce[2] = ce[3] = C(0);
  ce[4] = ce[5] = ce[6] = r3;
  ce[7] = ce[8] = ce[9] = C(0);
  ce[10] = -rx2 * uy;  // 210
  ce[11] = -rx2 * uz;  // 201
  ce[12] = -ry2 * ux;  // 120
  ce[13] = -ry2 * uz;  // 021
  ce[14] = -rz2 * ux;  // 102
  ce[15] = -rz2 * uy;  // 012
  ce[16] = -rxyz;      // 111
  ce[17] = C(3) * rx2 * uy * uy;  // 220
  ce[18] = C(3) * rx2 * uz * uz;  // 202
  ce[19] = C(3) * ry2 * uz * uz;  // 022
  ce[20] = C(3) * rxyz * ux;      // 211
  ce[21] = C(3) * rxyz * uy;      // 121
  ce[22] = C(3) * rxyz * uz;      // 112
  ce[23] = C(-6) * rxyz * ux * uy;  // 221
  ce[24] = C(-6) * rxyz * ux * uz;  // 212
  ce[25] = C(-6) * rxyz * uy * uz;  // 122
  ce[26] = C(0);                    // 222
}

// The equilibrium enters only through its cumulants, taken in closed form
// (eq_cumulants27) for density rho_eq.
template <typename C>
__device__ __forceinline__ void geier27(C (&f)[27], C rho_eq, const Macro<C>& mc, C omega, const C* rate) {
  using L = D3Q27;
  C m[27], kt[27], c[27], ce[27];
  moments<L>(f, m);
  to_central<L>(m, mc.ux, mc.uy, mc.uz);
  eq_cumulants27(rho_eq, mc.ux, mc.uy, mc.uz, ce);
  // C = rho * C~(kappa / rho), component 0 carries rho
  {
    const C r = m[0], ir = C(1) / r;
#pragma unroll
    for (int i = 0; i < 27; ++i) kt[i] = m[i] * ir;
    tables::central_to_cumulant27(kt, c);
#pragma unroll
    for (int i = 0; i < 27; ++i) c[i] *= r;
    c[0] = r;
  }
  const C ob = rate[0], oo = rate[1], oe = rate[2];
  const C d4 = c[4] - ce[4], d5 = c[5] - ce[5], d6 = c[6] - ce[6];
  const C tr3 = (d4 + d5 + d6) / C(3);
  c[4] = c[4] - (omega * (d4 - tr3) + ob * tr3);
  c[5] = c[5] - (omega * (d5 - tr3) + ob * tr3);
  c[6] = c[6] - (omega * (d6 - tr3) + ob * tr3);
#pragma unroll
  for (int i = 7; i < 10; ++i) c[i] = c[i] - omega * (c[i] - ce[i]);
#pragma unroll
  for (int i = 10; i < 17; ++i) c[i] = c[i] - oo * (c[i] - ce[i]);
#pragma unroll
  for (int i = 17; i < 27; ++i) c[i] = c[i] - oe * (c[i] - ce[i]);
  {
    const C r = c[0], ir = C(1) / r;
#pragma unroll
    for (int i = 0; i < 27; ++i) ce[i] = c[i] * ir;  // reuse ce as C~
    tables::cumulant_to_central27(ce, kt);
#pragma unroll
    for (int i = 0; i < 27; ++i) m[i] = kt[i] * r;
    m[0] = r;
  }
  to_raw<L>(m, mc.ux, mc.uy, mc.uz);
  populations<L>(m, f);
}

// Legacy D3Q27 "cumulant" of collide_cumulant_d3q27 (cumulant.py:935-1047):
// central moments of f_neq pass through the pseudo-cumulant map of
// cumulant.py:470-577 (Geier's partition terms without 1/rho; orders <= 3
// unchanged), relax (trace / deviatoric at omega_b / omega, 3rd order at
// omega_odd, 4th-6th at omega_even), map back (cumulant.py:580-651), unshift.
// rate = {omega_b, omega_odd, omega_even}.  Index order: DEG27.
template <typename C>
__device__ __forceinline__ void legacy_to_pseudo27(C (&k)[27]) {
  const C a = k[4], b = k[5], c = k[6], xy = k[7], xz = k[8], yz = k[9];
  k[17] = k[17] - a * b - C(2) * xy * xy;
  k[18] = k[18] - a * c - C(2) * xz * xz;
  k[19] = k[19] - b * c - C(2) * yz * yz;
  k[20] = k[20] - a * yz - C(2) * xz * xy;
  k[21] = k[21] - b * xz - C(2) * xy * yz;
  k[22] = k[22] - c * xy - C(2) * xz * yz;
  k[23] = k[23] - a * k[13] - C(2) * xy * k[16];
  k[24] = k[24] - a * k[15] - C(2) * xz * k[16];
  k[25] = k[25] - b * k[14] - C(2) * yz * k[16];
  // the 6th-order line reads the already mapped 4th-order entries
  k[26] = k[26] - a * k[19] - b * k[18] - c * k[17] - C(2) * (xy * k[22] + xz * k[21] + yz * k[20]) +
          C(2) * a * b * c + C(4) * xy * xz * yz + C(2) * xy * xy * c + C(2) * xz * xz * b + C(2) * yz * yz * a;
}

template <typename C>
__device__ __forceinline__ void legacy_from_pseudo27(C (&k)[27]) {
  const C a = k[4], b = k[5], c = k[6], xy = k[7], xz = k[8], yz = k[9];
  // 6th order first: it reads the mapped (pseudo-cumulant) 4th-order entries
  k[26] = k[26] + a * k[19] + b * k[18] + c * k[17] + C(2) * (xy * k[22] + xz * k[21] + yz * k[20]) -
          C(2) * a * b * c - C(4) * xy * xz * yz - C(2) * xy * xy * c - C(2) * xz * xz * b - C(2) * yz * yz * a;
  k[17] = k[17] + a * b + C(2) * xy * xy;
  k[18] = k[18] + a * c + C(2) * xz * xz;
  k[19] = k[19] + b * c + C(2) * yz * yz;
  k[20] = k[20] + a * yz + C(2) * xz * xy;
  k[21] = k[21] + b * xz + C(2) * xy * yz;
  k[22] = k[22] + c * xy + C(2) * xz * yz;
  k[23] = k[23] + a * k[13] + C(2) * xy * k[16];
  k[24] = k[24] + a * k[15] + C(2) * xz * k[16];
  k[25] = k[25] + b * k[14] + C(2) * yz * k[16];
}

template <typename C>
__device__ __forceinline__ void legacy27(C (&f)[27], const C (&feq)[27], const Macro<C>& mc, C omega,
                                         const C* rate) {
  using L = D3Q27;
  C v[27], m[27];
#pragma unroll
  for (int q = 0; q < 27; ++q) v[q] = f[q] - feq[q];
  moments<L>(v, m);
  to_central<L>(m, mc.ux, mc.uy, mc.uz);
  legacy_to_pseudo27(m);
  const C tr = m[4] + m[5] + m[6];
  const C trs = (C(1) - rate[0]) * tr;
  const C third = C(1.0 / 3.0);
#pragma unroll
  for (int i = 4; i < 7; ++i) m[i] = (C(1) - omega) * (m[i] - tr * third) + trs * third;
#pragma unroll
  for (int i = 7; i < 10; ++i) m[i] = (C(1) - omega) * m[i];
#pragma unroll
  for (int i = 10; i < 17; ++i) m[i] = (C(1) - rate[1]) * m[i];
#pragma unroll
  for (int i = 17; i < 27; ++i) m[i] = (C(1) - rate[2]) * m[i];
  legacy_from_pseudo27(m);
  to_raw<L>(m, mc.ux, mc.uy, mc.uz);
  populations<L>(m, v);
#pragma unroll
  for (int q = 0; q < 27; ++q) f[q] = feq[q] + v[q];
}

}  // namespace tlbm
