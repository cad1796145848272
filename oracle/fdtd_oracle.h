/*
 * CPU oracle for the 3D Yee FDTD time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY. GPU results are compared against this library: it is loaded by the
 * test-suite, by the package smoke check and by the CPU-baseline leg of the
 * benchmark, never by the product library or the C++ host shim.
 *
 * Layout everywhere: structure-of-arrays, one array per component, linear
 * index idx = i + j*nx + k*nx*ny (x fastest), as in gDiamond/gdiamond.hpp:1197.
 * Component order: Ex Ey Ez Hx Hy Hz (0..5).
 * Coefficient order (full material): Cax Cbx Cay Cby Caz Cbz Dax Dbx Day Dby Daz Dbz.
 */
#ifndef FDTD_ORACLE_H
#define FDTD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Point source: value table[n] enters the update of `comp` at step n as
 * "- table[n]*dx" inside the curl bracket (J for E comps, M for H comps). */
typedef struct or_source {
  int comp;
  int64_t i, j, k;
  const float* table; /* indexed by absolute step */
} or_source;

/* DFT monitor state, reference convention kernels.cuh:87-113:
 *   re[c][f][p] += E(p) * cos(2 pi f n) * scale
 *   im[c][f][p] -= E(p) * sin(2 pi f n) * scale
 * accumulated after step n with the field value after that step. */
typedef struct or_dft {
  int ncomp;
  int comps[6];
  int axis;         /* 0: yz plane at i=index, 1: xz plane at j=index, 2: xy plane at k=index */
  int64_t index;
  int nf;
  const double* f_cps; /* frequencies in cycles per step */
  double scale;
  double* re;       /* [ncomp][nf][plane] */
  double* im;
} or_dft;

/* Material: full reference arrays (12) or isotropic lossless (Cb per cell,
 * Ca = Da = 1, Db scalar). */
typedef struct or_material {
  int full;
  const float* c[12];
  const float* cb; /* iso */
  float db;        /* iso */
} or_material;

/* FP32 sequential stepping, operand order of gdiamond.hpp:1193-1221
 * (E over the interior k-j-i, then H). Advances fields by nsteps starting at
 * absolute step n0. nthreads > 1 selects the OpenMP twin of
 * gdiamond_omp.hpp:130-159 (collapse(2) over k,j). */
int or_run_f32(int64_t nx, int64_t ny, int64_t nz, float dx,
               float* const fields[6], const or_material* mat,
               int nsrc, const or_source* src,
               int64_t n0, int64_t nsteps, or_dft* dft, int nthreads);

/* FP64 twin: identical loops with double fields and double-converted
 * coefficients / sources. */
int or_run_f64(int64_t nx, int64_t ny, int64_t nz, double dx,
               double* const fields[6], const or_material* mat,
               int nsrc, const or_source* src,
               int64_t n0, int64_t nsteps, or_dft* dft, int nthreads);

/* FP16-storage twin: every component c is stored as binary16 of
 * value * 2^-scale_exp[c]; values are decoded exactly to float, computed in
 * float with the operand order of or_run_f32 (isotropic Cb given in float)
 * and rounded to binary16 (nearest even) when they are stored:
 *   block == 0: after every half step (E, then H), as an untiled update that
 *               writes E and H to memory each step does; the DFT sees the
 *               stored values;
 *   block >= 1: after every `block` steps (the last block of the call may be
 *               shorter), as a temporally blocked update that keeps the
 *               intermediate steps in float registers; the DFT sees the float
 *               value of every step (before the block-end rounding).
 * Fields hold binary16 bit patterns. Isotropic material only. `threads`
 * host threads run the block mode (its float steps and the conversions);
 * every value is computed by the same operations whatever the count. */
int or_run_f16(int64_t nx, int64_t ny, int64_t nz, float dx, uint16_t* const fields[6],
               const int scale_exp[6], const float* cb, float db,
               int nsrc, const or_source* src, int64_t n0, int64_t nsteps, or_dft* dft,
               int64_t block, int threads);

/* The same for a 16-bit format fmt: 0 = binary16 (== or_run_f16), 1 =
 * bfloat16 (the upper half of a float, rounded to nearest even). */
int or_run_h16(int64_t nx, int64_t ny, int64_t nz, float dx, uint16_t* const fields[6],
               const int scale_exp[6], const float* cb, float db,
               int nsrc, const or_source* src, int64_t n0, int64_t nsteps, or_dft* dft,
               int64_t block, int threads, int fmt);

/* bfloat16 conversions used by or_run_h16 (round to nearest even on the bit
 * pattern, NaN kept quiet). */
uint16_t or_f32_to_bf16(float f);
float or_bf16_to_f32(uint16_t h);
void or_f32_to_bf16_n(const float* in, uint16_t* out, int64_t n);
void or_bf16_to_f32_n(const uint16_t* in, float* out, int64_t n);

/* binary16 conversions used by or_run_f16 (round to nearest even); the
 * array forms split long arrays over the OpenMP threads. */
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);
void or_f32_to_f16_n(const float* in, uint16_t* out, int64_t n);
void or_f16_to_f32_n(const uint16_t* in, float* out, int64_t n);

/* Literal form of the reference loop with full J/M arrays (6 arrays, zero
 * except where set per step), gdiamond.hpp:1164-1221. Used to pin the
 * point-source shortcut of or_run_f32. */
int or_run_f32_jm_arrays(int64_t nx, int64_t ny, int64_t nz, float dx,
                         float* const fields[6], const float* const coeff[12],
                         int nsrc, const or_source* src, int64_t n0, int64_t nsteps);

/* Restatement of gDiamond::gDiamond (gdiamond.hpp:20-73) +
 * set_FDTD_matrices_3D_structure (utils.h:542-690) with the default mask
 * (all false): writes the 12 coefficient arrays. */
void or_default_problem_coefficients(int64_t nx, int64_t ny, int64_t nz, float* const coeff[12]);

/* The reference default source, gdiamond.hpp:1189-1191:
 * Mz(t) = M_source_amp * sin(SOURCE_OMEGA * t * dt), float arithmetic. */
float or_default_mz(int64_t t);
/* Flat index of the default source cell, gdiamond.hpp:50. */
int64_t or_default_source_idx(int64_t nx, int64_t ny, int64_t nz);

/* DFT weights exactly as accumulated: w_c = cos(2 pi f n) * scale,
 * w_s = sin(2 pi f n) * scale, computed in double. */
void or_dft_weights(double f_cps, int64_t n, double scale, double* wc, double* ws);

#ifdef __cplusplus
}
#endif
#endif
