/*
 * gd_capi.h -- C ABI of libgdiamond_hip.so, the MI355X (gfx950) 3D Yee FDTD
 * time-stepping library.
 *
 * The reference exposes its GPU path only as members of the header-only class
 * gdiamond::gDiamond (gDiamond/gdiamond.hpp:17-177); every member allocates,
 * runs and frees its own device buffers. This ABI replaces those members with
 * a context whose device state persists across calls. The C++ class with the
 * reference's names is include/gdiamond/gdiamond.hpp, built on this ABI.
 *
 * Conventions
 *  - Host arrays are structure-of-arrays, one float array per component,
 *    linear index idx = i + j*nx + k*nx*ny (x fastest), as the reference
 *    (gdiamond.hpp:1197). The caller owns host buffers; the library copies.
 *  - Component ids: 0..5 = Ex Ey Ez Hx Hy Hz.
 *  - Coefficient order: Cax Cbx Cay Cby Caz Cbz Dax Dbx Day Dby Daz Dbz
 *    (gdiamond.hpp:870-890). Cb/Db already contain 1/dx (utils.h:656-686).
 *  - Every call returns int status: GD_OK (0) or a negative GD_ERR_*;
 *    gd_last_error(ctx) gives the text. No call exits the process (the
 *    reference's CUDACHECK calls exit(1), gdiamond_gpu.cuh:11-20).
 *  - One context per host thread; calls on one context are serialised. Each
 *    context owns one HIP stream of default priority on the device that was
 *    current at gd_create.
 *  - Only interior cells 1..n-2 on every axis are updated; boundary cells keep
 *    their values (the reference's PEC walls, gdiamond.hpp:1194-1221).
 */
#ifndef GDIAMOND_GD_CAPI_H
#define GDIAMOND_GD_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gd_ctx gd_ctx;

enum {
  GD_OK = 0,
  GD_ERR_ARG = -1,         /* bad argument (size, index, null pointer, range) */
  GD_ERR_STATE = -2,       /* call not valid in the current state (e.g. no material) */
  GD_ERR_HIP = -3,         /* HIP runtime error (text in gd_last_error) */
  GD_ERR_NOMEM = -4,       /* device or host allocation failed */
  GD_ERR_UNSUPPORTED = -5, /* unsupported configuration */
  GD_ERR_NODEVICE = -6     /* no HIP device visible */
};

enum { GD_EX = 0, GD_EY = 1, GD_EZ = 2, GD_HX = 3, GD_HY = 4, GD_HZ = 5 };

/* Stepping kernels.
 * GD_VARIANT_PLAIN: two launches per step (E kernel, then H kernel), fields
 *   updated in place. Replaces updateE/H_3Dmap_fix (kernels.cuh:337-394) and
 *   updateE/H_2Dmap (kernels.cuh:207-264) behind
 *   update_FDTD_gpu_3D_warp_underutilization_fix / update_FDTD_gpu_2D
 *   (gdiamond_gpu.cuh:809, :483).
 * GD_VARIANT_TILED: temporally tiled; one launch advances `blt` steps over
 *   overlapped xy tiles streamed along z, ping-ponging between two field
 *   copies. Replaces the mix-mapping kernels ver2/ver3
 *   (kernels_mm_ver2.cuh:488-720, kernels_mm_ver3.cuh:417-647) behind
 *   update_FDTD_mix_mapping_gpu_ver2/ver3 (gdiamond_gpu_mm_ver2.cuh:837,
 *   gdiamond_gpu_mm_ver3.cuh:478). Results are bit-identical to PLAIN.
 * GD_VARIANT_FUSED: one launch per step that updates E and then H in one
 *   pass along z, reading step n from one field copy and writing step n+1
 *   to the other (52 bytes per cell-update for FP32 against PLAIN's 76;
 *   twice the field memory). The same update as PLAIN
 *   (update_FDTD_gpu_3D_warp_underutilization_fix, gdiamond_gpu.cuh:896-913),
 *   bit for bit, in every storage format (16-bit formats round every step,
 *   as PLAIN does). */
enum { GD_VARIANT_PLAIN = 0, GD_VARIANT_TILED = 1, GD_VARIANT_FUSED = 2 };

/* Tiling of GD_VARIANT_TILED. Zero fields select the planner's choice.
 *  blt    : time steps per launch (1..4)
 *  rows   : y rows held by each thread
 *  waves  : wave64 wavefronts per workgroup (tile height = rows*waves)
 *  zchunk : z planes owned by one workgroup */
typedef struct gd_tiling {
  int blt;
  int rows;
  int waves;
  int zchunk;
} gd_tiling;

#define GD_MAX_BLT 8
#define GD_MAX_SOURCES 8

/* Field storage formats (gd_set_storage). */
enum {
  GD_STORAGE_F32 = 0, /* float fields, float arithmetic: the reference's format */
  GD_STORAGE_F64 = 1, /* double fields and arithmetic (PLAIN and FUSED): reference runs */
  GD_STORAGE_F16 = 2, /* binary16 of value*2^-k per component; float arithmetic */
  GD_STORAGE_BF16 = 3 /* bfloat16 of value*2^-k per component; float arithmetic */
};

/* ---- lifecycle --------------------------------------------------------- */

/* Create a context for an nx*ny*nz grid (each >= 3) with cell size dx (used
 * only in the source term "- J*dx", gdiamond.hpp:1200-1219). Fields start at 0.
 * Replaces gDiamond::gDiamond(Nx,Ny,Nz) (gdiamond.hpp:20) for device state. */
int gd_create(gd_ctx** out, size_t nx, size_t ny, size_t nz, float dx);
int gd_destroy(gd_ctx* ctx);
const char* gd_last_error(const gd_ctx* ctx);
/* Library version string, e.g. "gdiamond-hip 0.1 gfx950". */
const char* gd_version(void);

/* ---- storage ------------------------------------------------------------- */

/* Storage format of the fields (and of the isotropic coefficient). Must be set
 * before the material; resets the fields to zero and the step counter.
 * GD_STORAGE_F16 stores each component c as binary16 of value * 2^-scale_exp[c]
 * (scale_exp may be NULL: all 0). Values are decoded exactly to float,
 * computed in float with the reference operand order and rounded to binary16
 * (nearest even) when written to device memory: after each half step (E, then
 * H) by GD_VARIANT_PLAIN, once per launch of `blt` steps by GD_VARIANT_TILED,
 * which keeps the intermediate steps in FP32 registers (so TILED results
 * depend on the tiling's blt, and a run's last launch covers the remaining
 * nsteps % blt steps). The DFT accumulates the stored values (PLAIN) or the
 * FP32 value of every step (TILED). It halves the field bytes per cell-update
 * (24 B instead of 48 B; the material adds 1 B indexed, 2 B dense). No counterpart in the
 * reference, which stores FP32 only (gdiamond.hpp:799-891).
 * GD_STORAGE_BF16 follows the same rules with bfloat16 (the upper half of a
 * float: 8 significand bits against binary16's 11, float's exponent range)
 * rounded to nearest even. */
int gd_set_storage(gd_ctx* ctx, int storage, const int* scale_exp);
int gd_get_storage(const gd_ctx* ctx, int* storage, int* scale_exp);

/* ---- material ------------------------------------------------------------ */

/* Device formats of the isotropic material (gd_set_material_format), and what
 * gd_get_material_format reports. */
enum {
  GD_MATERIAL_AUTO = 0,    /* INDEXED when Cb has at most GD_MAX_MATERIALS distinct
                              values and the storage is not FP64, else DENSE */
  GD_MATERIAL_DENSE = 1,   /* one Cb per cell in FP32 (FP64 with GD_STORAGE_F64) */
  GD_MATERIAL_INDEXED = 2, /* one 1-byte material id per cell + an FP32 table of Cb values */
  GD_MATERIAL_FULL = 3     /* reported only: the 12-array form of gd_set_coefficients */
};
#define GD_MAX_MATERIALS 256

/* Choose how the isotropic Cb is held on the device; must precede the
 * material (default GD_MATERIAL_AUTO). Both forms keep Cb exact and give
 * bit-identical results: DENSE reads 4 B of material per cell, INDEXED 1 B
 * (and a table lookup per level), which runs faster (512^3 FP32 TILED: 220
 * against 211 Gcell/s) and holds a quarter of the memory. INDEXED is not
 * available with GD_STORAGE_F64. No counterpart in the reference, which
 * streams 12 FP32 coefficient arrays (gdiamond_gpu.cuh:871-887). */
int gd_set_material_format(gd_ctx* ctx, int format);
/* Format in use (DENSE, INDEXED or FULL; -1 before any material) and the
 * number of table entries of the INDEXED form (0 otherwise). */
int gd_get_material_format(const gd_ctx* ctx, int* format, int* ntable);

/* Full reference coefficient arrays (12 arrays of nx*ny*nz floats), the layout
 * every reference driver copies to the device (gdiamond_gpu.cuh:871-887). */
int gd_set_coefficients(gd_ctx* ctx, const float* const coeff[12]);
/* Isotropic lossless form: Ca = Da = 1, one Cb array shared by Ex/Ey/Ez, one
 * scalar Db. Bit-identical to gd_set_coefficients with those values. */
int gd_set_isotropic(gd_ctx* ctx, const float* cb, float db);
/* Isotropic form from relative permittivity: cb = dt/(eps_r*dx),
 * db = dt/dx, each evaluated in double and rounded to float once. */
int gd_set_isotropic_eps(gd_ctx* ctx, const float* eps_r, float dt);
/* Isotropic form by material id: cell q has Cb = table[id[q]] (ntable in
 * 1..GD_MAX_MATERIALS, every id < ntable). Bit-identical to gd_set_isotropic
 * with the expanded array. */
int gd_set_isotropic_indexed(gd_ctx* ctx, const uint8_t* id, const float* table, int ntable,
                             float db);

/* ---- sources --------------------------------------------------------------- */

/* Point source on `component` at cell (i,j,k): at step n the update of that
 * component subtracts table[n]*dx inside the curl bracket, exactly where the
 * reference subtracts J[idx]*dx / M[idx]*dx (gdiamond.hpp:1200-1219). The
 * table covers absolute steps 0..nsteps-1; a run past its end is an error.
 * At most GD_MAX_SOURCES sources. Replaces the per-step source copies of
 * gdiamond_gpu.cuh:568-570 and :167-170. */
int gd_add_source(gd_ctx* ctx, int component, size_t i, size_t j, size_t k,
                  const float* table, size_t nsteps);
int gd_clear_sources(gd_ctx* ctx);

/* ---- fields ---------------------------------------------------------------- */

int gd_set_fields(gd_ctx* ctx, const float* const fields[6]);
/* Field and material copies of 8 MiB and more per array go through two
 * 64 MiB pinned host buffers that the context allocates on its first such
 * copy and gd_destroy frees (the reference's drivers copy the coefficients in
 * and the fields back on every call, gdiamond_gpu.cuh:878-889, :921-926). */
int gd_get_fields(gd_ctx* ctx, float* const fields[6]);
/* Exact field values in double for every storage format. */
int gd_get_fields_f64(gd_ctx* ctx, double* const fields[6]);
int gd_get_field_f64(gd_ctx* ctx, int component, double* out);
/* Copy one field component, or one xy plane (k) of it, to the host. */
int gd_get_field(gd_ctx* ctx, int component, float* out);
int gd_get_plane_xy(gd_ctx* ctx, int component, size_t k, float* out);
/* Zero all fields and DFT sums; the step counter returns to 0. */
int gd_reset(gd_ctx* ctx);
/* Absolute index of the next step to run. */
int gd_get_step(const gd_ctx* ctx, size_t* step);

/* ---- DFT monitor ----------------------------------------------------------- */

/* Running DFT of the components in comp_mask (bit c = component c) on one
 * plane: axis 2 -> xy plane k=index (update_field_FFT_xy, kernels.cuh:87),
 * axis 0 -> yz plane i=index (update_field_FFT_yz, kernels.cuh:60), axis 1 ->
 * xz plane j=index. After each step n, with v the component value after step
 * n, for every frequency f (cycles per step):
 *     re += v * (cos(2*pi*f*n) * scale);  im -= v * (sin(2*pi*f*n) * scale)
 * in FP64 (the reference uses phase omega*t with t = n*dt, so
 * f = freq*dt, and FP32 sums). Plane index: axis 2 -> i + j*nx, axis 0 ->
 * j + k*ny, axis 1 -> i + k*nx. Reconfiguring clears the sums. */
int gd_dft_config(gd_ctx* ctx, unsigned comp_mask, int axis, size_t index,
                  const double* f_cycles_per_step, size_t nf, double scale);
/* re/im: nf*plane doubles each for `component` (must be in the mask). */
int gd_dft_get(gd_ctx* ctx, int component, double* re, double* im);
int gd_dft_disable(gd_ctx* ctx);

/* ---- stepping ---------------------------------------------------------------- */

/* Advance nsteps steps from the current step with the given kernel variant
 * (tiling may be NULL). Synchronous: returns after the device finished. With
 * the planner's 4 x 8 tile a TILED launch runs the tiles that hold a source
 * cell or monitored cells on a second stream of the context, alongside the
 * others; both streams have the default priority. TILED and FUSED need an
 * xy plane below 512 Mi cells (GD_ERR_UNSUPPORTED otherwise; PLAIN has no
 * such limit). */
int gd_run(gd_ctx* ctx, int variant, size_t nsteps, const gd_tiling* tiling);

/* The same run, enqueued on the context's stream and returned from at once,
 * so the caller's host work overlaps it (the reference's drivers always wait,
 * cudaDeviceSynchronize before the copies back, gdiamond_gpu.cuh:914).
 * gd_sync waits for it; so does the next call on the context that changes it
 * or reads its results (fields, DFT sums, gd_last_run_stats), in stream order,
 * so a result read after gd_run_async is the run's result. Argument errors
 * are returned at once; errors of the device work surface at gd_sync. */
int gd_run_async(gd_ctx* ctx, int variant, size_t nsteps, const gd_tiling* tiling);
int gd_sync(gd_ctx* ctx);
/* The tiling the planner would use for GD_VARIANT_TILED on this context. */
int gd_plan_tiling(const gd_ctx* ctx, gd_tiling* out);
/* The planner's tiling for a grid and storage format, without a context or
 * a device.
 * Planner rule: 4 rows x 8 waves; 3 levels for 16-bit fields, 4 otherwise;
 * z chunks sized for ~1024 workgroups within 16..32 planes (16..64 for 16-bit
 * fields). No counterpart in the reference,
 * whose tile sizes are compile-time constants (kernels_mm_ver2.cuh:8-32). */
int gd_plan_tiling_for(size_t nx, size_t ny, size_t nz, int storage, gd_tiling* out);

/* ---- one TILED launch that allocates nothing ---------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one TILED launch of nsteps (1..GD_NOALLOC_MAX_STEPS)
 * steps on caller-owned device buffers, with FP32 fields and the isotropic
 * material (dense Cb, or 1-byte ids and a table of GD_MAX_MATERIALS floats).
 * It enqueues on `stream` and returns without waiting; it calls no hipMalloc,
 * hipFree or other allocating routine. Its scratch memory (source
 * descriptors, DFT weights, the per-level snapshot of the monitored plane) is
 * one fixed device array of 4 KiB + 4 * GD_NOALLOC_MAX_SNAP_CELLS floats
 * (32 MiB) that each GPU keeps until the process ends; every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams or
 * threads) race. The first call on a device looks up the array's address:
 * make it before capturing a graph. Sizes beyond the limits (among them an
 * xy plane of 512 Mi cells or more) are rejected with GD_ERR_ARG (text in
 * gd_last_error(NULL)). fields_out receives every cell
 * (boundary cells copied) and must not overlap fields_in; the DFT sums are
 * accumulated in place with the weights of absolute steps n0..n0+nsteps-1,
 * exactly as gd_run does, so a sequence of launches with n0 advancing by
 * nsteps and the field buffers swapped equals gd_run bit for bit. */
#define GD_NOALLOC_MAX_STEPS 4
#define GD_NOALLOC_MAX_FREQS 16
#define GD_NOALLOC_MAX_SNAP_CELLS (1 << 21) /* monitored components x plane cells */

typedef struct gd_point_source {
  int comp;           /* component 0..5 */
  int i, j, k;        /* cell */
  const float* table; /* device array: the value entering step n is table[n] */
} gd_point_source;

typedef struct gd_noalloc_launch {
  int nx, ny, nz;
  float dx;
  const float* fields_in[6]; /* device: Ex Ey Ez Hx Hy Hz before step n0 */
  float* fields_out[6];      /* device: the fields after step n0 + nsteps - 1 */
  const float* cb;           /* device: Cb per cell, used when material_id is NULL */
  const uint8_t* material_id;   /* device: material id per cell, or NULL */
  const float* material_table;  /* device: Cb of each id, GD_MAX_MATERIALS floats */
  float db;                  /* Db (H coefficient, scalar) */
  int nsrc;                  /* 0..GD_MAX_SOURCES point sources */
  const gd_point_source* sources; /* host array of nsrc descriptors */
  int n0;                    /* absolute step of the first of the nsteps steps */
  int nsteps;                /* 1..GD_NOALLOC_MAX_STEPS */
  unsigned dft_mask;         /* monitored components (bit c); 0: no monitor */
  int dft_axis, dft_index;   /* plane, as gd_dft_config */
  int nf;                    /* 1..GD_NOALLOC_MAX_FREQS */
  const double* f_cycles_per_step; /* host: nf frequencies */
  double dft_scale;
  double* dft_re;            /* device: [popcount(mask)][nf][plane] FP64 sums */
  double* dft_im;
  void* stream;              /* hipStream_t; NULL is the null stream */
} gd_noalloc_launch;

int gd_tiled_launch_noalloc(const gd_noalloc_launch* launch);

/* Timing of the last gd_run, from HIP events on the context's stream:
 * total_ms spans the whole stepping loop; with per-launch timing enabled
 * (gd_set_launch_timing) kernel_ms is the sum of the durations of the
 * stepping-kernel launches and nlaunch their number. */
int gd_last_run_stats(const gd_ctx* ctx, double* total_ms, double* kernel_ms, size_t* nlaunch);
int gd_set_launch_timing(gd_ctx* ctx, int enable);

/* Bytes of device memory held by the context. */
int gd_device_bytes(const gd_ctx* ctx, size_t* bytes);

/* Plain device-to-device copy kernel over n floats of two scratch buffers the
 * context allocates; used to calibrate HBM traffic counters (bytes read and
 * written per launch are exactly 4*n each). Returns kernel time in ms. */
int gd_calibration_copy(gd_ctx* ctx, size_t n, int reps, double* ms);

/* Number of visible HIP devices (0 when none). */
int gd_device_count(int* count);

/* ---- batches of independent problems -------------------------------------- */

/* A batch holds many independent FP32 problems of (possibly) different sizes
 * that advance together. Instances with equal dims form a size class; each
 * class has a tiling (blt, rows x waves, zchunk) chosen by gd_batch_autotune
 * or set explicitly. Classes whose tilings share (blt, rows, waves) run in
 * one persistent launch per blt steps: its workgroups take (instance, tile)
 * work items from a queue, largest instances first. Launch groups with
 * different shapes run concurrently as parallel branches.
 * Per instance the results equal those of a gd_ctx with the same problem run
 * with GD_VARIANT_TILED (and so the sequential update), bit for bit.
 * No counterpart in the reference, which runs one problem per process and
 * fixes its tile shapes with compile-time macros (kernels_mm_ver2.cuh:8-32). */
typedef struct gd_batch gd_batch;

/* Run modes of gd_batch_run. */
enum {
  GD_BATCH_PERSISTENT = 0,  /* persistent launches issued directly */
  GD_BATCH_GRAPH = 1,       /* the same launches captured once per nsteps in a
                               hipGraph, replayed on later runs */
  GD_BATCH_PER_INSTANCE = 2 /* one ordinary TILED launch sequence per instance
                               (the unbatched way; for comparison) */
};

/* dims: 3*ninst sizes (nx, ny, nz per instance, each >= 3, at most 2^20,
 * an xy plane below 512 Mi cells, at most 2^36 cells); every instance has
 * cell size dx. Fields start at zero. */
int gd_batch_create(gd_batch** out, int ninst, const size_t* dims, float dx);
int gd_batch_destroy(gd_batch* b);
const char* gd_batch_last_error(const gd_batch* b);
/* Isotropic material of instance `inst` (as gd_set_isotropic / _eps). Launches
 * use 1-byte material ids when every instance's Cb has at most
 * GD_MAX_MATERIALS distinct values, the dense FP32 Cb otherwise (same
 * results). */
int gd_batch_set_isotropic(gd_batch* b, int inst, const float* cb, float db);
int gd_batch_set_isotropic_eps(gd_batch* b, int inst, const float* eps_r, float dt);
/* Point source of instance `inst` (as gd_add_source). */
int gd_batch_add_source(gd_batch* b, int inst, int comp, size_t i, size_t j, size_t k,
                        const float* table, size_t nsteps);
int gd_batch_set_fields(gd_batch* b, int inst, const float* const fields[6]);
int gd_batch_get_fields(gd_batch* b, int inst, float* const fields[6]);
/* Fields of every instance to zero, step counter to 0. */
int gd_batch_reset(gd_batch* b);
int gd_batch_get_step(const gd_batch* b, size_t* step);
/* Size classes (distinct dims, in order of first appearance). */
int gd_batch_num_classes(const gd_batch* b, int* nclass);
int gd_batch_class_info(const gd_batch* b, int cls, size_t dims[3], int* ninst);
int gd_batch_set_class_tiling(gd_batch* b, int cls, const gd_tiling* t);
int gd_batch_get_class_tiling(const gd_batch* b, int cls, gd_tiling* t);
/* Autotuner: for each size class, times every supported tiling (blt 1..4,
 * rows x waves in {2,4}x{4,8}, zchunk in {16,32,64,nz}) as persistent launches
 * over that class's instances and keeps, among the tilings within 3 % of the
 * fastest per cell-update, the one with the most levels per launch, then the
 * tallest tile, then the fastest. Fields
 * and the step counter are left unchanged (the timing launches always read
 * the current copy). `reps` timed launches per candidate (>= 1). */
int gd_batch_autotune(gd_batch* b, int reps);
/* Advance every instance by nsteps (all source tables must cover them). Each
 * persistent launch of B steps runs the items whose tile holds a source cell
 * with the full kernel, then the other items with a build without source
 * code, on the group's stream. */
int gd_batch_run(gd_batch* b, size_t nsteps, int mode);
/* Wall time of the last gd_batch_run on the device (events around the whole
 * run) and the number of kernel launches it made. */
int gd_batch_last_run_stats(const gd_batch* b, double* total_ms, size_t* nlaunch);
int gd_batch_device_bytes(const gd_batch* b, size_t* bytes);

/* ---- learned coarse-grid corrector: network kernels --------------------------- */

/* A 3-layer 3x3x3 convolution network (12 -> 16 -> 16 -> 12 channels,
 * LeakyReLU(1/8) after layers 1 and 2, zero padding) on one nx*ny*nz grid,
 * BF16 activations and weights, FP32 accumulation (MFMA), FP32 master
 * parameters with Adam. Parameters: 3 blocks of W[27][16][16] (tap, output
 * channel, input channel; tap = (dz+1)*9 + (dy+1)*3 + (dx+1)) followed by
 * b[16], FP32; channels beyond 12 on the input of layer 1 and the output of
 * layer 3 are padding (held at zero). Tensors are channels-last, 16 FP32
 * values per cell, cell = x + nx*(y + ny*z). No counterpart in the reference,
 * which has no learned component. */
typedef struct gd_nn gd_nn;
int gd_nn_create(gd_nn** out, size_t nx, size_t ny, size_t nz);
int gd_nn_destroy(gd_nn* nn);
const char* gd_nn_last_error(const gd_nn* nn);
int gd_nn_num_params(size_t* n);
/* Sets the parameters and resets the Adam moments. */
int gd_nn_set_params(gd_nn* nn, const float* params, size_t n);
int gd_nn_get_params(gd_nn* nn, float* params, size_t n);
/* Input features [cell][16] (rounded to BF16 on upload). */
int gd_nn_set_input(gd_nn* nn, const float* x);
/* Training target [cell][16] (channels 0..11 used) and loss mask per cell. */
int gd_nn_set_target(gd_nn* nn, const float* target, const uint8_t* mask);
/* Forward pass; y_out (optional) receives the FP32 output [cell][16]. */
int gd_nn_forward(gd_nn* nn, float* y_out);
/* Forward, MSE over masked cells x 12 channels, backward, Adam step with
 * learning rate lr (lr <= 0: gradients only). loss: before the update;
 * grads_out (optional): the parameter gradients. */
int gd_nn_train_step(gd_nn* nn, float lr, float* loss, float* grads_out);
/* Restriction of the current fields of a context (cell size h, FP32) to the
 * grid of cell size 2h with nx/2 x ny/2 x nz/2 cells, consistent with the Yee
 * positions (Ex(i+1/2,j,k), Hx(i,j+1/2,k+1/2), ...): each coarse E component
 * is the mean of the two fine edges along its axis, each H component the mean
 * of the four fine faces of its face. out: 6 host arrays of the coarse size. */
int gd_restrict_fields(gd_ctx* ctx, float* const out[6]);
/* GD_VARIANT_TILED run with the corrector applied after every `interval`
 * steps (the last step of each interval runs on its own so that E one step
 * back is known): features from E, H, E - E_prev and the material, network
 * forward, F += y / r on interior cells. cfg = {se, sh, sd, re, rh, dx/dt}:
 * feature scales of E, H, E - E_prev, output scales of E and H, and dx/dt of
 * the problem (1/eps = cb*dx/dt). Steps past the last full interval run
 * uncorrected. FP32 fields, isotropic material (dense or indexed), network
 * on the same grid. */
int gd_run_corrected(gd_ctx* ctx, gd_nn* nn, size_t nsteps, int interval, const float cfg[6],
                     const gd_tiling* tiling);
/* D = A B for 16x16 matrices (row-major floats, rounded to BF16) computed
 * with the MFMA operand layout the convolutions use. */
int gd_nn_mfma_selftest(const float* a, const float* b, float* d);

/* ---- host-side problem builders (no device needed) -------------------------- */

/* The reference default problem: gDiamond::gDiamond coefficients
 * (gdiamond.hpp:20-73 + set_FDTD_matrices_3D_structure, utils.h:542-690;
 * all-false mask, graded absorber, dx = 0.1, dt = 0.05). */
int gd_default_problem_coefficients(size_t nx, size_t ny, size_t nz, float* const coeff[12]);
/* Reference default source values: Mz(t) = M_amp*sin(omega*t*dt) in float
 * (gdiamond.hpp:1189-1191), t = 0..nsteps-1. */
int gd_default_mz_table(size_t nsteps, float* out);

/* Seeded random sphere medium: eps_r = 1 background; nspheres spheres with
 * centres uniform in [0,n) per axis, radius U[rmin,rmax], eps U[emin,emax];
 * later spheres overwrite earlier ones; a cell is inside when
 * |(i,j,k)-c|^2 <= r^2. RNG: SplitMix64, u = (x>>11) * 2^-53, draws in the
 * order cx, cy, cz, r, eps per sphere. */
int gd_sphere_medium(size_t nx, size_t ny, size_t nz, uint64_t seed, int nspheres,
                     double rmin, double rmax, double emin, double emax, float* eps_out);

/* Figures. gd_field_rgb: the reference's field colormap (save_field_png,
 * utils.h:210-269): value = u/vmax clamped to [-1,1], white -> red for
 * positive, white -> blue for negative values, pixel (i,j) = u[i + j*nx];
 * rgb: 3*nx*ny bytes, row j. gd_save_field_png writes that image as a PNG
 * (nx wide, ny high); gd_save_mask_png writes black where mask != 0, white
 * elsewhere (save_mask_png, utils.h:170-207); gd_write_png_rgb writes any
 * 8-bit RGB image. The PNG encoder is the library's own (stored deflate
 * blocks): the files differ from the reference's stb output, the decoded
 * pixels are the same. */
int gd_field_rgb(const float* u, int nx, int ny, float vmax, uint8_t* rgb);
int gd_save_field_png(const float* u, const char* filename, int nx, int ny, float vmax);
int gd_save_mask_png(const uint8_t* mask, const char* filename, int nx, int ny);
int gd_write_png_rgb(const char* filename, int w, int h, const uint8_t* rgb);

/* Gaussian-modulated sine J(n) = exp(-((n-t0)/tau)^2) * sin(2*pi*f*n),
 * evaluated in double, rounded to float once, n = 0..nsteps-1. */
int gd_gaussian_pulse_table(size_t nsteps, double t0, double tau, double f_cps, float* out);

#ifdef __cplusplus
}
#endif
#endif /* GDIAMOND_GD_CAPI_H */
