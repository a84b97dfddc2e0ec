// Velocity sets as compile-time tables.  Orderings are the reference's:
//   D2Q9  src/tensorlbm/d2q9.py:7-27   (carried in 3-D with cz = 0)
//   D3Q19 src/tensorlbm/d3q19.py:6-67  (OPPOSITE[q] != 18 - q)
//   D3Q27 src/tensorlbm/d3q27.py:32-84
// Weights are exact rationals; kernels read them as compile-time constants.
#pragma once

namespace tlbm {

struct D2Q9 {
  static constexpr int Q = 9;
  static constexpr int cx[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int cz[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr double w[Q] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                                  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
};

struct D3Q19 {
  static constexpr int Q = 19;
  static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
  static constexpr double w[Q] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                                  1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
};

struct D3Q27 {
  static constexpr int Q = 27;
  static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1,
                                -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
  static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0,
                                0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1,
                                -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
  static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12,
                                 11, 18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};
  static constexpr double w[Q] = {
      8.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,
      1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,
      1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 216, 1.0 / 216,
      1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216};
};

// Compile-time self checks of the tables: weights sum to one, opposites are
// involutive and negate the velocity.
template <class L>
constexpr bool lattice_ok() {
  double s = 0.0;
  for (int q = 0; q < L::Q; ++q) {
    s += L::w[q];
    const int o = L::opp[q];
    if (L::opp[o] != q) return false;
    if (L::cx[o] != -L::cx[q] || L::cy[o] != -L::cy[q] || L::cz[o] != -L::cz[q]) return false;
    if (L::w[o] != L::w[q]) return false;
  }
  return s > 1.0 - 1e-15 && s < 1.0 + 1e-15;
}
static_assert(lattice_ok<D2Q9>(), "D2Q9 table");
static_assert(lattice_ok<D3Q19>(), "D3Q19 table");
static_assert(lattice_ok<D3Q27>(), "D3Q27 table");

}  // namespace tlbm
