/*
 * C restatement of the reference's fused periodic step, TEST INFRASTRUCTURE
 * (oracle): out = collide_bgk(stream(f)) with the fused-kernel moment
 * convention of src/tensorlbm/triton_fused.py:396-436 (pull stream with
 * periodic wrap, rho floored at 1e-12 inside f_eq, omega = 1/tau).
 * Optional full-way bounce-back at solid cells (triton_fused_obstacle.py:
 * 1168-1222): the solid cell takes the pulled population of the opposite
 * lane.  OpenMP over z-y rows.  Used by the tests as a second CPU checker
 * and timed by bench.py as the CPU baseline.
 */
#include <stdint.h>
#include <stddef.h>

#define DEFINE_STEP(NAME, T)                                                                       \
  int NAME(int Q, const int* c, const int* opp, const double* w, const T* fin, T* fout,            \
           const int8_t* solid, int nz, int ny, int nx, double tau) {                              \
    if (Q > 27 || Q < 1) return -1;                                                                \
    const size_t N = (size_t)nz * ny * nx;                                                         \
    const T omega = (T)(1.0 / tau);                                                                \
    _Pragma("omp parallel for collapse(2) schedule(static)")                                      \
    for (int z = 0; z < nz; ++z) {                                                                 \
      for (int y = 0; y < ny; ++y) {                                                               \
        T f[27], cxq[27], cyq[27], czq[27], wq[27];                                                \
        for (int q = 0; q < Q; ++q) {                                                              \
          cxq[q] = (T)c[3 * q]; cyq[q] = (T)c[3 * q + 1]; czq[q] = (T)c[3 * q + 2]; wq[q] = (T)w[q]; \
        }                                                                                          \
        for (int x = 0; x < nx; ++x) {                                                             \
          const size_t cell = ((size_t)z * ny + y) * nx + x;                                       \
          for (int q = 0; q < Q; ++q) {                                                            \
            int sx = x - c[3 * q], sy = y - c[3 * q + 1], sz = z - c[3 * q + 2];                   \
            sx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);                                     \
            sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);                                     \
            sz = sz < 0 ? sz + nz : (sz >= nz ? sz - nz : sz);                                     \
            f[q] = fin[(size_t)q * N + ((size_t)sz * ny + sy) * nx + sx];                          \
          }                                                                                        \
          if (solid && solid[cell]) {                                                              \
            T g[27];                                                                               \
            for (int q = 0; q < Q; ++q) g[q] = f[opp[q]];                                          \
            for (int q = 0; q < Q; ++q) f[q] = g[q];                                               \
          }                                                                                        \
          T rho = 0, jx = 0, jy = 0, jz = 0;                                                       \
          for (int q = 0; q < Q; ++q) {                                                            \
            rho += f[q]; jx += cxq[q] * f[q]; jy += cyq[q] * f[q]; jz += czq[q] * f[q];            \
          }                                                                                        \
          const T rs = rho > (T)1e-12 ? rho : (T)1e-12;                                            \
          const T ux = jx / rs, uy = jy / rs, uz = jz / rs;                                        \
          const T usq = ux * ux + uy * uy + uz * uz;                                               \
          for (int q = 0; q < Q; ++q) {                                                            \
            const T cu = cxq[q] * ux + cyq[q] * uy + czq[q] * uz;                                  \
            const T feq = rs * wq[q] * ((T)1 + (T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * usq);       \
            fout[(size_t)q * N + cell] = f[q] - omega * (f[q] - feq);                              \
          }                                                                                        \
        }                                                                                          \
      }                                                                                            \
    }                                                                                              \
    return 0;                                                                                      \
  }

DEFINE_STEP(oracle_fused_bgk_f32, float)
DEFINE_STEP(oracle_fused_bgk_f64, double)

int oracle_num_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}
