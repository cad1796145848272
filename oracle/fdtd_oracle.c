/*
 * CPU oracle for the 3D Yee FDTD time-stepping path -- TEST INFRASTRUCTURE.
 *
 * Only the test-suite, the package smoke check and the benchmark's CPU-baseline
 * leg load this library, always as the reference results are compared
 * against. The product never links it.
 *
 * Every function restates the reference algorithm it names; pinning of each
 * function (golden vectors from the reference build, closed forms, invariants)
 * lives in tests/test_oracle_*.py.
 *
 * Build: gcc -O3 -fopenmp -ffp-contract=off (x86-64 baseline: no FMA), so each
 * float operation rounds exactly once, as the reference CPU build does
 * (CMakeLists.txt:23) and as its CUDA build forces with --fmad=false
 * (CMakeLists.txt:25).
 */
#include "fdtd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI_F 3.14159265359f /* utils.h:17 */
#define OR_PI_D 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* point sources                                                       */
/* ------------------------------------------------------------------ */

/* The reference keeps full J/M arrays that are zero except at the source
 * (gdiamond.hpp:1184-1191) and always subtracts J[idx]*dx. For every cell
 * without a source that term is (+0)*dx = +0 and x - (+0) == x bit for bit,
 * so applying the term only at source cells is exact. or_run_f32_jm_arrays
 * keeps the literal array form to pin this. */
static inline float src_val_f(int nsrc, const or_source* src, int comp,
                              int64_t i, int64_t j, int64_t k, int64_t n) {
  float v = 0.0f;
  for (int s = 0; s < nsrc; ++s)
    if (src[s].comp == comp && src[s].i == i && src[s].j == j && src[s].k == k)
      v = src[s].table[n];
  return v;
}

static int64_t plane_size(const or_dft* d, int64_t nx, int64_t ny, int64_t nz) {
  if (d->axis == 0) return ny * nz;
  if (d->axis == 1) return nx * nz;
  return nx * ny;
}

/* cos and sin are called separately and never fused into sincos(): glibc's
 * sincos() differs from sin()/cos() in the last bit for some arguments. */
static __attribute__((noinline)) double or_cos(double a) { return cos(a); }
static __attribute__((noinline)) double or_sin(double a) { return sin(a); }

void or_dft_weights(double f_cps, int64_t n, double scale, double* wc, double* ws) {
  const double ang = 2.0 * OR_PI_D * f_cps * (double)n;
  *wc = or_cos(ang) * scale;
  *ws = or_sin(ang) * scale;
}

/* kernels.cuh:87-113 convention, FP64 accumulators. */
#define DEFINE_DFT(NAME, T)                                                             \
  static void NAME(or_dft* d, int64_t nx, int64_t ny, int64_t nz, T* const f[6],        \
                   int64_t n) {                                                          \
    if (!d || d->ncomp <= 0) return;                                                     \
    const int64_t P = plane_size(d, nx, ny, nz);                                         \
    const int64_t sxy = nx * ny;                                                         \
    for (int c = 0; c < d->ncomp; ++c) {                                                 \
      const T* a = f[d->comps[c]];                                                       \
      for (int q = 0; q < d->nf; ++q) {                                                  \
        double wc, ws;                                                                   \
        or_dft_weights(d->f_cps[q], n, d->scale, &wc, &ws);                              \
        double* re = d->re + ((int64_t)c * d->nf + q) * P;                               \
        double* im = d->im + ((int64_t)c * d->nf + q) * P;                               \
        if (d->axis == 2) {                                                              \
          const T* pl = a + d->index * sxy;                                              \
          for (int64_t p = 0; p < P; ++p) {                                              \
            re[p] += (double)pl[p] * wc;                                                 \
            im[p] -= (double)pl[p] * ws;                                                 \
          }                                                                              \
        } else if (d->axis == 1) {                                                       \
          for (int64_t k = 0; k < nz; ++k)                                               \
            for (int64_t i = 0; i < nx; ++i) {                                           \
              const double e = (double)a[i + d->index * nx + k * sxy];                   \
              re[i + k * nx] += e * wc;                                                  \
              im[i + k * nx] -= e * ws;                                                  \
            }                                                                            \
        } else {                                                                         \
          for (int64_t k = 0; k < nz; ++k)                                               \
            for (int64_t j = 0; j < ny; ++j) {                                           \
              const double e = (double)a[d->index + j * nx + k * sxy];                   \
              re[j + k * ny] += e * wc;                                                  \
              im[j + k * ny] -= e * ws;                                                  \
            }                                                                            \
        }                                                                                \
      }                                                                                  \
    }                                                                                    \
  }

DEFINE_DFT(dft_acc_f32, float)
DEFINE_DFT(dft_acc_f64, double)

/* ------------------------------------------------------------------ */
/* stepping, gdiamond.hpp:1164-1221 (seq) / gdiamond_omp.hpp:110-159   */
/* ------------------------------------------------------------------ */

/* One body generator for float and double. Operand order is the
 * reference's: Ca*E + Cb*(((a - b) - (c - d)) - J*dx). */
#define DEFINE_RUN(NAME, T, DFTACC)                                                      \
  int NAME(int64_t nx, int64_t ny, int64_t nz, T dx, T* const F[6],                      \
           const or_material* mat, int nsrc, const or_source* src, int64_t n0,           \
           int64_t nsteps, or_dft* dft, int nthreads) {                                  \
    if (nx < 3 || ny < 3 || nz < 3) return -1;                                           \
    T* Ex = F[0]; T* Ey = F[1]; T* Ez = F[2];                                            \
    T* Hx = F[3]; T* Hy = F[4]; T* Hz = F[5];                                            \
    const int64_t sxy = nx * ny;                                                         \
    const float* const* C = mat->c;                                                      \
    const int full = mat->full;                                                          \
    const float* cb = mat->cb;                                                           \
    const T db = (T)mat->db;                                                             \
    int nth = nthreads < 1 ? 1 : nthreads;                                               \
    for (int64_t n = n0; n < n0 + nsteps; ++n) {                                         \
      /* update E */                                                                     \
      _Pragma("omp parallel for collapse(2) schedule(static) num_threads(nth) if(nth > 1)") \
      for (int64_t k = 1; k < nz - 1; k++) {                                             \
        for (int64_t j = 1; j < ny - 1; j++) {                                           \
          for (int64_t i = 1; i < nx - 1; i++) {                                         \
            const int64_t idx = i + j * nx + k * sxy;                                    \
            T jx = 0, jy = 0, jz = 0;                                                    \
            if (nsrc) {                                                                  \
              jx = (T)src_val_f(nsrc, src, 0, i, j, k, n);                               \
              jy = (T)src_val_f(nsrc, src, 1, i, j, k, n);                               \
              jz = (T)src_val_f(nsrc, src, 2, i, j, k, n);                               \
            }                                                                            \
            const T cx = ((Hz[idx] - Hz[idx - nx]) - (Hy[idx] - Hy[idx - sxy])) - jx * dx; \
            const T cy = ((Hx[idx] - Hx[idx - sxy]) - (Hz[idx] - Hz[idx - 1])) - jy * dx;  \
            const T cz = ((Hy[idx] - Hy[idx - 1]) - (Hx[idx] - Hx[idx - nx])) - jz * dx;   \
            if (full) {                                                                  \
              Ex[idx] = (T)C[0][idx] * Ex[idx] + (T)C[1][idx] * cx;                      \
              Ey[idx] = (T)C[2][idx] * Ey[idx] + (T)C[3][idx] * cy;                      \
              Ez[idx] = (T)C[4][idx] * Ez[idx] + (T)C[5][idx] * cz;                      \
            } else {                                                                     \
              const T b = (T)cb[idx];                                                    \
              Ex[idx] = Ex[idx] + b * cx;                                                \
              Ey[idx] = Ey[idx] + b * cy;                                                \
              Ez[idx] = Ez[idx] + b * cz;                                                \
            }                                                                            \
          }                                                                              \
        }                                                                                \
      }                                                                                  \
      /* update H */                                                                     \
      _Pragma("omp parallel for collapse(2) schedule(static) num_threads(nth) if(nth > 1)") \
      for (int64_t k = 1; k < nz - 1; k++) {                                             \
        for (int64_t j = 1; j < ny - 1; j++) {                                           \
          for (int64_t i = 1; i < nx - 1; i++) {                                         \
            const int64_t idx = i + j * nx + k * sxy;                                    \
            T mx = 0, my = 0, mz = 0;                                                    \
            if (nsrc) {                                                                  \
              mx = (T)src_val_f(nsrc, src, 3, i, j, k, n);                               \
              my = (T)src_val_f(nsrc, src, 4, i, j, k, n);                               \
              mz = (T)src_val_f(nsrc, src, 5, i, j, k, n);                               \
            }                                                                            \
            const T cx = ((Ey[idx + sxy] - Ey[idx]) - (Ez[idx + nx] - Ez[idx])) - mx * dx; \
            const T cy = ((Ez[idx + 1] - Ez[idx]) - (Ex[idx + sxy] - Ex[idx])) - my * dx;  \
            const T cz = ((Ex[idx + nx] - Ex[idx]) - (Ey[idx + 1] - Ey[idx])) - mz * dx;   \
            if (full) {                                                                  \
              Hx[idx] = (T)C[6][idx] * Hx[idx] + (T)C[7][idx] * cx;                      \
              Hy[idx] = (T)C[8][idx] * Hy[idx] + (T)C[9][idx] * cy;                      \
              Hz[idx] = (T)C[10][idx] * Hz[idx] + (T)C[11][idx] * cz;                    \
            } else {                                                                     \
              Hx[idx] = Hx[idx] + db * cx;                                               \
              Hy[idx] = Hy[idx] + db * cy;                                               \
              Hz[idx] = Hz[idx] + db * cz;                                               \
            }                                                                            \
          }                                                                              \
        }                                                                                \
      }                                                                                  \
      DFTACC(dft, nx, ny, nz, F, n);                                                     \
    }                                                                                    \
    return 0;                                                                            \
  }

DEFINE_RUN(or_run_f32, float, dft_acc_f32)
DEFINE_RUN(or_run_f64, double, dft_acc_f64)

/* ------------------------------------------------------------------ */
/* binary16 storage                                                     */
/* ------------------------------------------------------------------ */

/* Nearest binary16 by exact arithmetic: the spacing of binary16 numbers near
 * |f| is q = 2^(e-11) for |f| >= 2^-14 (f = m*2^e, m in [0.5,1)) and 2^-24
 * below; |f|/q is exact and rint() rounds it to nearest even. */
uint16_t or_f32_to_f16(float f) {
  const uint16_t sign = signbit(f) ? 0x8000u : 0u;
  if (isnan(f)) return (uint16_t)(sign | 0x7e00u);
  const double a = fabs((double)f);
  if (a >= 65520.0) return (uint16_t)(sign | 0x7c00u);
  if (a == 0.0) return sign;
  int e;
  frexp(a, &e);
  const int qe = a >= ldexp(1.0, -14) ? e - 11 : -24;
  const double n = rint(ldexp(a, -qe)); /* integer count of quanta */
  if (qe == -24) return (uint16_t)(sign | (uint16_t)n); /* subnormal (1024 = min normal) */
  int E = e - 1 + 15;
  int mant = (int)n - 1024;
  if ((int)n == 2048) { E += 1; mant = 0; }
  if (E >= 31) return (uint16_t)(sign | 0x7c00u);
  return (uint16_t)(sign | (E << 10) | mant);
}

float or_f16_to_f32(uint16_t h) {
  const int E = (h >> 10) & 31, mant = h & 1023;
  double v;
  if (E == 31) v = mant ? NAN : INFINITY;
  else if (E == 0) v = ldexp((double)mant, -24);
  else v = ldexp((double)(mant + 1024), E - 25);
  return (float)((h & 0x8000u) ? -v : v);
}

/* bfloat16 = the upper 16 bits of a float, rounded to nearest even on the bit
 * pattern (exact integer arithmetic; NaN stays a quiet NaN). */
uint16_t or_f32_to_bf16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(x >> 16);
}

float or_bf16_to_f32(uint16_t h) {
  const uint32_t x = (uint32_t)h << 16;
  float f;
  memcpy(&f, &x, 4);
  return f;
}

void or_f32_to_bf16_n(const float* in, uint16_t* out, int64_t n) {
#pragma omp parallel for schedule(static) if (n > (1 << 22))
  for (int64_t q = 0; q < n; ++q) out[q] = or_f32_to_bf16(in[q]);
}

void or_bf16_to_f32_n(const uint16_t* in, float* out, int64_t n) {
#pragma omp parallel for schedule(static) if (n > (1 << 22))
  for (int64_t q = 0; q < n; ++q) out[q] = or_bf16_to_f32(in[q]);
}

void or_f32_to_f16_n(const float* in, uint16_t* out, int64_t n) {
#pragma omp parallel for schedule(static) if (n > (1 << 22))
  for (int64_t q = 0; q < n; ++q) out[q] = or_f32_to_f16(in[q]);
}

void or_f16_to_f32_n(const uint16_t* in, float* out, int64_t n) {
#pragma omp parallel for schedule(static) if (n > (1 << 22))
  for (int64_t q = 0; q < n; ++q) out[q] = or_f16_to_f32(in[q]);
}

/* 16-bit storage element <-> float: fmt 0 = binary16, 1 = bfloat16. */
static inline float dec16(int fmt, uint16_t h, float s) {
  return (fmt ? or_bf16_to_f32(h) : or_f16_to_f32(h)) * s;
}
static inline uint16_t enc16(int fmt, float v, float inv) {
  return fmt ? or_f32_to_bf16(v * inv) : or_f32_to_f16(v * inv);
}

static void dft_acc_f16(int fmt, or_dft* d, int64_t nx, int64_t ny, int64_t nz,
                        uint16_t* const f[6], const float* sc, int64_t n) {
  if (!d || d->ncomp <= 0) return;
  const int64_t P = plane_size(d, nx, ny, nz), sxy = nx * ny;
  for (int c = 0; c < d->ncomp; ++c) {
    const int comp = d->comps[c];
    const uint16_t* a = f[comp];
    for (int q = 0; q < d->nf; ++q) {
      double wc, ws;
      or_dft_weights(d->f_cps[q], n, d->scale, &wc, &ws);
      double* re = d->re + ((int64_t)c * d->nf + q) * P;
      double* im = d->im + ((int64_t)c * d->nf + q) * P;
      for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
          for (int64_t i = 0; i < nx; ++i) {
            int64_t p;
            if (d->axis == 2) { if (k != d->index) continue; p = i + j * nx; }
            else if (d->axis == 1) { if (j != d->index) continue; p = i + k * nx; }
            else { if (i != d->index) continue; p = j + k * ny; }
            const double e = (double)dec16(fmt, a[i + j * nx + k * sxy], sc[comp]);
            re[p] += e * wc;
            im[p] -= e * ws;
          }
    }
  }
}

/* Block mode of or_run_f16: blocks of `block` steps (the last one shorter)
 * run in float from the decoded state, then the state is rounded to binary16. */
static int run_f16_blocks(int fmt, int64_t nx, int64_t ny, int64_t nz, float dx,
                          uint16_t* const F[6],
                          const float* sc, const float* inv, const float* cb, float db,
                          int nsrc, const or_source* src, int64_t n0, int64_t nsteps,
                          or_dft* dft, int64_t block, int nth) {
  const int64_t N = nx * ny * nz;
  float* W[6] = {0};
  int rc = 0;
  for (int c = 0; c < 6; ++c) {
    W[c] = (float*)malloc((size_t)N * sizeof(float));
    if (!W[c]) { rc = -2; goto done; }
    float* w = W[c];
    const uint16_t* f = F[c];
    const float s = sc[c];
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
    for (int64_t q = 0; q < N; ++q) w[q] = dec16(fmt, f[q], s);
  }
  or_material m;
  memset(&m, 0, sizeof(m));
  m.full = 0;
  m.cb = cb;
  m.db = db;
  for (int64_t n = n0; n < n0 + nsteps;) {
    const int64_t nb = n0 + nsteps - n < block ? n0 + nsteps - n : block;
    rc = or_run_f32(nx, ny, nz, dx, W, &m, nsrc, src, n, nb, dft, nth);
    if (rc) goto done;
    for (int c = 0; c < 6; ++c) {
      float* w = W[c];
      uint16_t* f = F[c];
      const float s = sc[c], iv = inv[c];
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
      for (int64_t q = 0; q < N; ++q) {
        f[q] = enc16(fmt, w[q], iv);
        w[q] = dec16(fmt, f[q], s);
      }
    }
    n += nb;
  }
done:
  for (int c = 0; c < 6; ++c) free(W[c]);
  return rc;
}

int or_run_h16(int64_t nx, int64_t ny, int64_t nz, float dx, uint16_t* const F[6],
               const int scale_exp[6], const float* cb, float db,
               int nsrc, const or_source* src, int64_t n0, int64_t nsteps, or_dft* dft,
               int64_t block, int threads, int fmt) {
  if (nx < 3 || ny < 3 || nz < 3 || block < 0 || fmt < 0 || fmt > 1) return -1;
  const int nth = threads > 1 ? threads : 1;
  float sc[6], inv[6];
  for (int c = 0; c < 6; ++c) {
    sc[c] = ldexpf(1.0f, scale_exp ? scale_exp[c] : 0);
    inv[c] = ldexpf(1.0f, scale_exp ? -scale_exp[c] : 0);
  }
  if (block > 0)
    return run_f16_blocks(fmt, nx, ny, nz, dx, F, sc, inv, cb, db, nsrc, src, n0, nsteps, dft,
                          block, nth);
  uint16_t *Ex = F[0], *Ey = F[1], *Ez = F[2], *Hx = F[3], *Hy = F[4], *Hz = F[5];
  const int64_t sxy = nx * ny;
  for (int64_t n = n0; n < n0 + nsteps; ++n) {
    for (int64_t k = 1; k < nz - 1; k++)
      for (int64_t j = 1; j < ny - 1; j++)
        for (int64_t i = 1; i < nx - 1; i++) {
          const int64_t idx = i + j * nx + k * sxy;
          const float hz = dec16(fmt, Hz[idx], sc[5]), hy = dec16(fmt, Hy[idx], sc[4]),
                      hx = dec16(fmt, Hx[idx], sc[3]);
          const float cx = ((hz - dec16(fmt, Hz[idx - nx], sc[5])) - (hy - dec16(fmt, Hy[idx - sxy], sc[4]))) -
                           src_val_f(nsrc, src, 0, i, j, k, n) * dx;
          const float cy = ((hx - dec16(fmt, Hx[idx - sxy], sc[3])) - (hz - dec16(fmt, Hz[idx - 1], sc[5]))) -
                           src_val_f(nsrc, src, 1, i, j, k, n) * dx;
          const float cz = ((hy - dec16(fmt, Hy[idx - 1], sc[4])) - (hx - dec16(fmt, Hx[idx - nx], sc[3]))) -
                           src_val_f(nsrc, src, 2, i, j, k, n) * dx;
          const float b = cb[idx];
          Ex[idx] = enc16(fmt, dec16(fmt, Ex[idx], sc[0]) + b * cx, inv[0]);
          Ey[idx] = enc16(fmt, dec16(fmt, Ey[idx], sc[1]) + b * cy, inv[1]);
          Ez[idx] = enc16(fmt, dec16(fmt, Ez[idx], sc[2]) + b * cz, inv[2]);
        }
    for (int64_t k = 1; k < nz - 1; k++)
      for (int64_t j = 1; j < ny - 1; j++)
        for (int64_t i = 1; i < nx - 1; i++) {
          const int64_t idx = i + j * nx + k * sxy;
          const float ex = dec16(fmt, Ex[idx], sc[0]), ey = dec16(fmt, Ey[idx], sc[1]),
                      ez = dec16(fmt, Ez[idx], sc[2]);
          const float cx = ((dec16(fmt, Ey[idx + sxy], sc[1]) - ey) - (dec16(fmt, Ez[idx + nx], sc[2]) - ez)) -
                           src_val_f(nsrc, src, 3, i, j, k, n) * dx;
          const float cy = ((dec16(fmt, Ez[idx + 1], sc[2]) - ez) - (dec16(fmt, Ex[idx + sxy], sc[0]) - ex)) -
                           src_val_f(nsrc, src, 4, i, j, k, n) * dx;
          const float cz = ((dec16(fmt, Ex[idx + nx], sc[0]) - ex) - (dec16(fmt, Ey[idx + 1], sc[1]) - ey)) -
                           src_val_f(nsrc, src, 5, i, j, k, n) * dx;
          Hx[idx] = enc16(fmt, dec16(fmt, Hx[idx], sc[3]) + db * cx, inv[3]);
          Hy[idx] = enc16(fmt, dec16(fmt, Hy[idx], sc[4]) + db * cy, inv[4]);
          Hz[idx] = enc16(fmt, dec16(fmt, Hz[idx], sc[5]) + db * cz, inv[5]);
        }
    dft_acc_f16(fmt, dft, nx, ny, nz, F, sc, n);
  }
  return 0;
}

int or_run_f16(int64_t nx, int64_t ny, int64_t nz, float dx, uint16_t* const F[6],
               const int scale_exp[6], const float* cb, float db,
               int nsrc, const or_source* src, int64_t n0, int64_t nsteps, or_dft* dft,
               int64_t block, int threads) {
  return or_run_h16(nx, ny, nz, dx, F, scale_exp, cb, db, nsrc, src, n0, nsteps, dft, block,
                    threads, 0);
}

/* Literal reference loop with J/M arrays (gdiamond.hpp:1164-1221): the arrays
 * are zero, each step writes the source values at their cells and every cell
 * subtracts J[idx]*dx. */
int or_run_f32_jm_arrays(int64_t nx, int64_t ny, int64_t nz, float dx,
                         float* const F[6], const float* const C[12],
                         int nsrc, const or_source* src, int64_t n0, int64_t nsteps) {
  const int64_t N = nx * ny * nz, sxy = nx * ny;
  float* JM[6];
  for (int c = 0; c < 6; ++c) {
    JM[c] = (float*)calloc((size_t)N, sizeof(float));
    if (!JM[c]) return -2;
  }
  float *Ex = F[0], *Ey = F[1], *Ez = F[2], *Hx = F[3], *Hy = F[4], *Hz = F[5];
  float *Jx = JM[0], *Jy = JM[1], *Jz = JM[2], *Mx = JM[3], *My = JM[4], *Mz = JM[5];
  for (int64_t t = n0; t < n0 + nsteps; t++) {
    for (int s = 0; s < nsrc; ++s)
      JM[src[s].comp][src[s].i + src[s].j * nx + src[s].k * sxy] = src[s].table[t];
    for (int64_t k = 1; k < nz - 1; k++)
      for (int64_t j = 1; j < ny - 1; j++)
        for (int64_t i = 1; i < nx - 1; i++) {
          int64_t idx = i + j * nx + k * sxy;
          Ex[idx] = C[0][idx] * Ex[idx] + C[1][idx] *
            ((Hz[idx] - Hz[idx - nx]) - (Hy[idx] - Hy[idx - sxy]) - Jx[idx] * dx);
          Ey[idx] = C[2][idx] * Ey[idx] + C[3][idx] *
            ((Hx[idx] - Hx[idx - sxy]) - (Hz[idx] - Hz[idx - 1]) - Jy[idx] * dx);
          Ez[idx] = C[4][idx] * Ez[idx] + C[5][idx] *
            ((Hy[idx] - Hy[idx - 1]) - (Hx[idx] - Hx[idx - nx]) - Jz[idx] * dx);
        }
    for (int64_t k = 1; k < nz - 1; k++)
      for (int64_t j = 1; j < ny - 1; j++)
        for (int64_t i = 1; i < nx - 1; i++) {
          int64_t idx = i + j * nx + k * sxy;
          Hx[idx] = C[6][idx] * Hx[idx] + C[7][idx] *
            ((Ey[idx + sxy] - Ey[idx]) - (Ez[idx + nx] - Ez[idx]) - Mx[idx] * dx);
          Hy[idx] = C[8][idx] * Hy[idx] + C[9][idx] *
            ((Ez[idx + 1] - Ez[idx]) - (Ex[idx + sxy] - Ex[idx]) - My[idx] * dx);
          Hz[idx] = C[10][idx] * Hz[idx] + C[11][idx] *
            ((Ex[idx + nx] - Ex[idx]) - (Ey[idx + 1] - Ey[idx]) - Mz[idx] * dx);
        }
  }
  for (int c = 0; c < 6; ++c) free(JM[c]);
  return 0;
}

/* ------------------------------------------------------------------ */
/* default problem, gdiamond.hpp:20-73 + utils.h:542-690               */
/* ------------------------------------------------------------------ */

/* Complex float arithmetic of utils.h:31-116 (each product/quotient in float). */
typedef struct { float re, im; } cpx;
static inline cpx cmul(cpx a, cpx b) {
  cpx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static inline cpx cdiv(cpx a, cpx b) {
  const float den = b.re * b.re + b.im * b.im;
  cpx r;
  r.re = (a.re * b.re + a.im * b.im) / den;
  r.im = (a.im * b.re - a.re * b.im) / den;
  return r;
}

/* Physical constants of the default problem. eps0 = mu0 = 1 (utils.h:19-22);
 * dx = wavelength/10 = 0.1f, dt = 0.05f, omega = 2*PI*c0/lambda in float
 * (gdiamond.hpp:894-906). */
static float def_dx(void) { return 1.0f / 10; }
static float def_dt(void) { return 0.05f; }
static float def_omega(void) {
  const float c0 = 1.0f / sqrtf(1.0f * 1.0f);
  const float freq = c0 / 1.0f;
  return 2 * OR_PI_F * freq;
}

/* Grading of one axis: distance into the absorbing layer (utils.h:609-618). */
static float bound_dist(int i, int N, int t_pml) {
  float d = 0;
  if (i < t_pml) d = (float)(1.0 - (double)((float)i / (float)t_pml));
  if (i + t_pml >= N) d = (float)(1.0 - (double)((float)(N - i - 1) / (float)t_pml));
  return d;
}

void or_default_problem_coefficients(int64_t nx64, int64_t ny64, int64_t nz64, float* const C[12]) {
  const int Nx = (int)nx64, Ny = (int)ny64, Nz = (int)nz64;
  const float dx = def_dx(), dt = def_dt(), omega0 = def_omega();
  const float eps0 = 1.0f, mu0 = 1.0f;
  const float eta0 = sqrtf(mu0 / eps0);
  /* t_PML = ceil(1 um / dx) (gdiamond.hpp:63-64) */
  const int t_pml = (int)ceilf(1.0f / dx);
  const float a_max = 2.0f;
  const int p = 3;
  const float den = 2 * eta0 * (float)t_pml * dx;
  const float sigma_max = (float)((double)(-(p + 1)) * log(1e-5) / (double)den);
  for (int i = 0; i < Nx; ++i)
    for (int j = 0; j < Ny; ++j)
      for (int k = 0; k < Nz; ++k) {
        /* the default mask is all-false: eps_r = eps_air = 1 + 0i everywhere */
        cpx eps_r = {1.0f, 0.0f};
        cpx mu_r = {1.0f, 0.0f};
        const int64_t idx = (int64_t)i + (int64_t)j * Nx + (int64_t)k * Nx * Ny;
        cpx s[3];
        const int ijk[3] = {i, j, k}, NN[3] = {Nx, Ny, Nz};
        for (int a = 0; a < 3; ++a) {
          const float d = bound_dist(ijk[a], NN[a], t_pml);
          const double pw = pow((double)d, (double)p);
          s[a].re = (float)(1 + (double)a_max * pw);
          s[a].im = (float)((double)sigma_max * pw / (double)(omega0 * eps0));
        }
        /* eps_xx = eps*sy*sz/sx etc. (utils.h:646-653) */
        cpx e[3], m[3];
        e[0] = cdiv(cmul(cmul(eps_r, s[1]), s[2]), s[0]);
        e[1] = cdiv(cmul(cmul(eps_r, s[0]), s[2]), s[1]);
        e[2] = cdiv(cmul(cmul(eps_r, s[0]), s[1]), s[2]);
        m[0] = cdiv(cmul(cmul(mu_r, s[1]), s[2]), s[0]);
        m[1] = cdiv(cmul(cmul(mu_r, s[0]), s[2]), s[1]);
        m[2] = cdiv(cmul(cmul(mu_r, s[0]), s[1]), s[2]);
        for (int a = 0; a < 3; ++a) {
          float er = e[a].re; if (er < 1) er = 1;
          float mr = m[a].re; if (mr < 1) mr = 1;
          const float se = fabsf(omega0 * eps0 * e[a].im);
          const float sh = fabsf(omega0 * mu0 * m[a].im);
          float tmp = (float)((double)(se * dt) / (2.0 * (double)er * (double)eps0));
          C[2 * a][idx] = (float)((1.0 - (double)tmp) / (1.0 + (double)tmp));
          C[2 * a + 1][idx] = (float)((double)(dt / (er * eps0)) / (1.0 + (double)tmp) / (double)dx);
          tmp = (float)((double)(sh * dt) / (2.0 * (double)mr * (double)mu0));
          C[6 + 2 * a][idx] = (float)((1.0 - (double)tmp) / (1.0 + (double)tmp));
          C[6 + 2 * a + 1][idx] = (float)((double)(dt / (mr * mu0)) / (1.0 + (double)tmp) / (double)dx);
        }
      }
}

float or_default_mz(int64_t t) {
  /* M_source_amp = J_source_amp * pow(eta0, 3.0) = 5e4 (float member) */
  const float m_amp = (float)(5e4f * pow((double)sqrtf(1.0f / 1.0f), 3.0));
  return m_amp * sinf(def_omega() * (float)t * def_dt());
}

int64_t or_default_source_idx(int64_t nx, int64_t ny, int64_t nz) {
  return nx / 2 + ny / 2 * nx + nz / 2 * nx * ny;
}
