// Learned coarse-grid corrector: a 3-layer 3x3x3 convolution network
// (12 -> 16 -> 16 -> 12 channels, LeakyReLU(1/8) between layers) on BF16
// activations and weights with FP32 accumulation on the MFMA units, with the
// kernels its training needs (backward data / weight passes, loss, Adam).
//
// No counterpart in the reference (it has no learned component); the fine
// grid solver stays the reference its output is measured against.
//
// Tensors are channels-last with 16 channels per cell, cell = x + nx*(y + ny*z)
// (the field layout): one 8-byte load gives a lane the 4 consecutive channels
// an MFMA operand needs. Convolutions are implicit GEMMs: a wave computes 16
// output channels x 16 consecutive x cells and accumulates 27 taps, each one
// v_mfma_f32_16x16x16bf16_1k with K = 16 input channels. Operand layout of
// that instruction (lane l, j = 0..3):
//   A (16x16, M x K): A[l % 16][4*(l/16) + j]
//   B (16x16, K x N): B[4*(l/16) + j][l % 16]
//   D (16x16 FP32):   D[4*(l/16) + j][l % 16]
// gd_nn_mfma_selftest checks this layout on the device against a host product.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gdiamond/gd_capi.h"
#include "gd_common.hpp"

namespace gd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kLeak = 0.125f;  // LeakyReLU slope (a power of two: exact)

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

namespace {

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct NnGrid {
  int nx, ny, nz;
  int nsegx;  // 16-cell x segments per row
};

// Implicit-GEMM 3x3x3 convolution with zero padding.
//   in:   bf16 [cell][16]
//   w:    bf16 [27][16 (output channel m)][16 (input channel k)], tap = (dz+1)*9+(dy+1)*3+(dx+1)
//   out:  kOut == 0: bf16 [cell][16] after bias and activation
//         kOut == 1: fp32 [cell][16] after bias (no activation)
//   kAct: 0 none, 1 LeakyReLU (forward), 2 multiply by LeakyReLU'(mask[cell][m])
//         (backward: mask holds the post-activation values of that layer)
// The input is staged in LDS: a workgroup of 4 waves
// owns a 16 x 4 x 4 output tile; its input with a one-cell halo
// (18 x 6 x 6 cells x 16 channels, 20.7 KB) is loaded once, every wave keeps
// the 27 weight operands of its lanes in registers and computes 4 rows of 16
// cells (27 MFMA each) from LDS.
constexpr int kTX = 16, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;

template <int kAct, int kOut>
__global__ __launch_bounds__(256) void conv3_lds_kernel(const uint16_t* __restrict__ in,
                                                        const uint16_t* __restrict__ w,
                                                        const float* __restrict__ bias,
                                                        void* __restrict__ out,
                                                        const uint16_t* __restrict__ mask,
                                                        NnGrid g) {
  __shared__ bf16x4 tile[kHZ * kHY * kHX * 4];
  const int ntx = (g.nx + kTX - 1) / kTX, nty = (g.ny + kTY - 1) / kTY;
  const int bx = blockIdx.x % ntx;
  const int by = (blockIdx.x / ntx) % nty;
  const int bz = blockIdx.x / (ntx * nty);
  const int x0 = bx * kTX, y0 = by * kTY, z0 = bz * kTZ;
  const long long sxy = (long long)g.nx * g.ny;
  for (int q = threadIdx.x; q < kHZ * kHY * kHX * 4; q += 256) {
    const int cq = q >> 2, ch = q & 3;
    const int hx = cq % kHX, hy = (cq / kHX) % kHY, hz = cq / (kHX * kHY);
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
    bf16x4 v = {0, 0, 0, 0};
    if (gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz)
      v = *reinterpret_cast<const bf16x4*>(
          in + ((long long)gx + (long long)gy * g.nx + (long long)gz * sxy) * 16 + 4 * ch);
    tile[q] = v;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  bf16x4 wr[27];
#pragma unroll
  for (int t = 0; t < 27; ++t)
    wr[t] = *reinterpret_cast<const bf16x4*>(w + ((long long)t * 16 + n) * 16 + 4 * kq);
  __syncthreads();
  const int x = x0 + n;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int row = wv * 4 + rr;  // 0..15
    const int ty = row % kTY, tz = row / kTY;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
      const int hc = ((tz + 1 + dz) * kHY + (ty + 1 + dy)) * kHX + (n + 1 + dx);
      acc = mfma16(wr[t], tile[hc * 4 + kq], acc);
    }
    const int y = y0 + ty, z = z0 + tz;
    if (x >= g.nx || y >= g.ny || z >= g.nz) continue;
    const long long c = (long long)x + (long long)y * g.nx + (long long)z * sxy;
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = 4 * kq + j;
      v[j] = acc[j] + (bias ? bias[m] : 0.0f);
      if (kAct == 1) v[j] = v[j] > 0.0f ? v[j] : v[j] * kLeak;
      if (kAct == 2) v[j] = bf16_to_f32(mask[c * 16 + m]) > 0.0f ? v[j] : v[j] * kLeak;
    }
    if (kOut == 1) {
      float* o = static_cast<float*>(out) + c * 16 + 4 * kq;
      *reinterpret_cast<float4*>(o) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
      bf16x4 r;
#pragma unroll
      for (int j = 0; j < 4; ++j) r[j] = (short)f32_to_bf16(v[j]);
      *reinterpret_cast<bf16x4*>(static_cast<uint16_t*>(out) + c * 16 + 4 * kq) = r;
    }
  }
}

// Weight gradient of one layer: dW[tap][m][k] = sum over cells p of
// dz[p][m] * act[p + tap][k], db[m] = sum dz[p][m]. Each wave reduces a range
// of `chunk` cells (a multiple of 16) into per-wave partials
// part[wave][27*256 + 16]; nn_reduce_kernel sums them in a fixed order.
__global__ __launch_bounds__(256) void conv3_wgrad_kernel(const uint16_t* __restrict__ dz,
                                                          const uint16_t* __restrict__ act,
                                                          float* __restrict__ part, NnGrid g,
                                                          long long ncell, int chunk) {
  const int lane = threadIdx.x & 63;
  const long long gw = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long p_begin = gw * chunk;
  if (p_begin >= ncell) return;  // whole wave
  const long long p_end = p_begin + chunk < ncell ? p_begin + chunk : ncell;
  const int col = lane & 15, kq = lane >> 4;
  const long long sxy = (long long)g.nx * g.ny;
  f32x4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc[t] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
  float dbias = 0.0f;
  for (long long p0 = p_begin; p0 < p_end; p0 += 16) {
    // this lane's 4 cells of the K = 16 cell group: p0 + 4*kq + j
    bf16x4 a;
    int cx[4], cy[4], cz[4];
    bool in_range[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long p = p0 + 4 * kq + j;
      in_range[j] = p < p_end;
      const long long pc = in_range[j] ? p : p_begin;
      cx[j] = (int)(pc % g.nx);
      cy[j] = (int)((pc / g.nx) % g.ny);
      cz[j] = (int)(pc / sxy);
      // A operand: A[m = col][k = cell j] = dz[cell][col]
      a[j] = in_range[j] ? (short)dz[pc * 16 + col] : (short)0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) dbias += bf16_to_f32((uint16_t)a[j]);
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int dzz = t / 9 - 1, dyy = (t / 3) % 3 - 1, dxx = t % 3 - 1;
      bf16x4 b;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int xx = cx[j] + dxx, yy = cy[j] + dyy, zz = cz[j] + dzz;
        const bool ok = in_range[j] && xx >= 0 && xx < g.nx && yy >= 0 && yy < g.ny && zz >= 0 &&
                        zz < g.nz;
        b[j] = ok ? (short)act[((long long)xx + (long long)yy * g.nx + (long long)zz * sxy) * 16 +
                              col]
                  : (short)0;
      }
      acc[t] = mfma16(a, b, acc[t]);
    }
  }
  // D[m = 4*kq + j][n = col]: dW[t][m][k = col]
  float* dst = part + gw * (27 * 256 + 16);
#pragma unroll
  for (int t = 0; t < 27; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[t * 256 + (4 * kq + j) * 16 + col] = acc[t][j];
  // dbias: lane (col, kq) summed dz[.][col] over its cells; reduce the 4 kq groups
  dbias += __shfl_xor(dbias, 16);
  dbias += __shfl_xor(dbias, 32);
  if (kq == 0) dst[27 * 256 + col] = dbias;
}

// out[q] = sum over w < nw of part[w][q] (fixed order), q < len.
__global__ __launch_bounds__(256) void nn_reduce_kernel(const float* __restrict__ part, int nw,
                                                        int len, float* __restrict__ out) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= len) return;
  float s = 0.0f;
  for (int w = 0; w < nw; ++w) s += part[(long long)w * len + q];
  out[q] = s;
}

// Mean squared error over the cells with mask[cell] != 0 and the first `nch`
// channels: dz = 2 (y - t) / count as bf16 [cell][16] (zero elsewhere),
// partial sums of (y - t)^2 per block in loss_part.
__global__ __launch_bounds__(256) void nn_loss_kernel(const float* __restrict__ y,
                                                      const float* __restrict__ t,
                                                      const uint8_t* __restrict__ mask,
                                                      long long ncell, int nch, float inv_count,
                                                      uint16_t* __restrict__ dz,
                                                      float* __restrict__ loss_part) {
  __shared__ float red[256];
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  float s = 0.0f;
  if (c < ncell) {
    const bool on = mask[c] != 0;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      float g = 0.0f;
      if (on && m < nch) {
        const float d = y[c * 16 + m] - t[c * 16 + m];
        s += d * d;
        g = 2.0f * d * inv_count;
      }
      dz[c * 16 + m] = f32_to_bf16(g);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Adam on FP32 master parameters, then the BF16 weight layouts for the
// forward pass (w[tap][m][k]) and the backward data pass (flipped taps,
// transposed channels: wt[26-tap][k][m]).
__global__ __launch_bounds__(256) void nn_adam_kernel(float* __restrict__ p,
                                                      const float* __restrict__ g,
                                                      float* __restrict__ m1,
                                                      float* __restrict__ m2, int n, float lr,
                                                      float b1, float b2, float eps, float c1,
                                                      float c2) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const float gq = g[q];
  const float a = b1 * m1[q] + (1.0f - b1) * gq;
  const float b = b2 * m2[q] + (1.0f - b2) * gq * gq;
  m1[q] = a;
  m2[q] = b;
  p[q] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// Input features of the corrector at one cell (channels-last, 16 values):
//   0-2  E * se            3-5  H * sh           6-8  (E - E_prev) * sd
//   9    (eps - 1) / 5     10   1/eps - 1        11   1 when a face neighbour
//                                                     has another eps, else 0
//   12-15 zero
// with 1/eps = cb * dx/dt of the isotropic material.
__global__ __launch_bounds__(256) void nn_features_kernel(Fields f, const float* __restrict__ pe0,
                                                          const float* __restrict__ pe1,
                                                          const float* __restrict__ pe2,
                                                          const float* __restrict__ cb,
                                                          NnFeat nf, int nx, int ny, int nz,
                                                          uint16_t* __restrict__ x) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long sxy = (long long)nx * ny;
  if (c >= sxy * nz) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sxy);
  const float* F[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) F[q] = static_cast<const float*>(f.c[q]);
  float v[16];
  v[0] = F[0][c] * nf.se;
  v[1] = F[1][c] * nf.se;
  v[2] = F[2][c] * nf.se;
  v[3] = F[3][c] * nf.sh;
  v[4] = F[4][c] * nf.sh;
  v[5] = F[5][c] * nf.sh;
  v[6] = (F[0][c] - pe0[c]) * nf.sd;
  v[7] = (F[1][c] - pe1[c]) * nf.sd;
  v[8] = (F[2][c] - pe2[c]) * nf.sd;
  const float cc = cb[c];
  const float inv_eps = cc * nf.dx_over_dt;
  v[9] = (1.0f / inv_eps - 1.0f) * 0.2f;
  v[10] = inv_eps - 1.0f;
  bool edge = false;
  if (i > 0) edge |= cb[c - 1] != cc;
  if (i + 1 < nx) edge |= cb[c + 1] != cc;
  if (j > 0) edge |= cb[c - nx] != cc;
  if (j + 1 < ny) edge |= cb[c + nx] != cc;
  if (k > 0) edge |= cb[c - sxy] != cc;
  if (k + 1 < nz) edge |= cb[c + sxy] != cc;
  v[11] = edge ? 1.0f : 0.0f;
#pragma unroll
  for (int q = 12; q < 16; ++q) v[q] = 0.0f;
#pragma unroll
  for (int q = 0; q < 16; q += 4) {
    bf16x4 r;
#pragma unroll
    for (int t = 0; t < 4; ++t) r[t] = (short)f32_to_bf16(v[q + t]);
    *reinterpret_cast<bf16x4*>(x + c * 16 + q) = r;
  }
}

// Correction of the fields from the network output: F_q += y[cell][q] / r_q
// (r = re for E, rh for H) on interior cells only (the walls stay frozen).
__global__ __launch_bounds__(256) void nn_apply_kernel(Fields f, const float* __restrict__ y,
                                                       NnFeat nf, int nx, int ny, int nz) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long sxy = (long long)nx * ny;
  if (c >= sxy * nz) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sxy);
  if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || k < 1 || k > nz - 2) return;
  const float ie = 1.0f / nf.re, ih = 1.0f / nf.rh;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    float* F = static_cast<float*>(f.c[q]);
    F[c] = F[c] + y[c * 16 + q] * (q < 3 ? ie : ih);
  }
}

// Dense Cb of an indexed material: out[q] = table[id[q]] (the corrector's
// features need Cb per cell).
__global__ __launch_bounds__(256) void expand_material_kernel(const uint8_t* __restrict__ id,
                                                              const float* __restrict__ table,
                                                              long long n,
                                                              float* __restrict__ out) {
  const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
  if (q < n) out[q] = table[id[q]];
}

// Yee-consistent restriction of fine fields (cell size h) to a grid of
// cell size 2h with cnx = nx/2, ... cells: E components average the two fine
// edges along their axis, H components the four fine faces of their face.
__global__ __launch_bounds__(256) void restrict_kernel(Fields f, int nx, int ny, int nz, int cnx,
                                                       int cny, int cnz, float* __restrict__ out) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long csxy = (long long)cnx * cny;
  const long long cn = csxy * cnz;
  if (c >= cn) return;
  const int I = (int)(c % cnx), J = (int)((c / cnx) % cny), K = (int)(c / csxy);
  const long long sxy = (long long)nx * ny;
  const long long o = 2ll * I + 2ll * J * nx + 2ll * K * sxy;
  const long long dxo = 1, dyo = nx, dzo = sxy;
  const float* F[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) F[q] = static_cast<const float*>(f.c[q]);
  out[0 * cn + c] = 0.5f * (F[0][o] + F[0][o + dxo]);
  out[1 * cn + c] = 0.5f * (F[1][o] + F[1][o + dyo]);
  out[2 * cn + c] = 0.5f * (F[2][o] + F[2][o + dzo]);
  out[3 * cn + c] = 0.25f * ((F[3][o] + F[3][o + dyo]) + (F[3][o + dzo] + F[3][o + dyo + dzo]));
  out[4 * cn + c] = 0.25f * ((F[4][o] + F[4][o + dxo]) + (F[4][o + dzo] + F[4][o + dxo + dzo]));
  out[5 * cn + c] = 0.25f * ((F[5][o] + F[5][o + dxo]) + (F[5][o + dyo] + F[5][o + dxo + dyo]));
}

// Parameter block of one layer: W[27][16][16] then b[16] (FP32, 6928 floats).
constexpr int kLayerParams = 27 * 256 + 16;

__global__ __launch_bounds__(256) void nn_pack_kernel(const float* __restrict__ p,
                                                      uint16_t* __restrict__ wf,
                                                      uint16_t* __restrict__ wb, int cin, int cout) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= 27 * 256) return;
  const int t = q / 256, m = (q / 16) % 16, k = q % 16;
  // padded channels stay zero
  const float v = (m < cout && k < cin) ? p[q] : 0.0f;
  const uint16_t h = f32_to_bf16(v);
  wf[q] = h;
  wb[(26 - t) * 256 + k * 16 + m] = h;
}

}  // namespace
}  // namespace gd

// ---- C ABI -------------------------------------------------------------------

struct gd_nn {
  int nx = 0, ny = 0, nz = 0;
  long long ncell = 0;
  // activations and gradients (bf16 [cell][16]) and the FP32 output
  uint16_t* x = nullptr;
  uint16_t* a1 = nullptr;
  uint16_t* a2 = nullptr;
  float* y = nullptr;
  float* target = nullptr;
  uint8_t* mask = nullptr;
  uint16_t* dz3 = nullptr;
  uint16_t* dz2 = nullptr;
  uint16_t* dz1 = nullptr;
  // parameters: 3 layers x kLayerParams FP32, Adam moments, gradients
  float* p = nullptr;
  float* g = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  uint16_t* wf = nullptr;  // [3][27*256] forward layouts
  uint16_t* wb = nullptr;  // [3][27*256] backward-data layouts
  float* part = nullptr;   // weight-gradient partials
  int nwg = 0;             // waves of the weight-gradient pass
  int chunk = 0;
  float* loss_part = nullptr;
  int nloss = 0;
  long long adam_t = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return GD_ERR_HIP;
  }
};

#define NH(o, call)                                        \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return (o)->hip_fail(e_, #call); \
  } while (0)

namespace {

thread_local std::string g_nn_create_error;

constexpr int kNParams = 3 * gd::kLayerParams;
const int kCin[3] = {12, 16, 16};
const int kCout[3] = {16, 16, 12};

gd::NnGrid nn_grid(const gd_nn* o) {
  gd::NnGrid g;
  g.nx = o->nx;
  g.ny = o->ny;
  g.nz = o->nz;
  g.nsegx = (o->nx + 15) / 16;
  return g;
}

unsigned conv_blocks(const gd_nn* o) {
  const long long t = (long long)((o->nx + gd::kTX - 1) / gd::kTX) *
                      ((o->ny + gd::kTY - 1) / gd::kTY) * ((o->nz + gd::kTZ - 1) / gd::kTZ);
  return (unsigned)t;
}

int pack_weights(gd_nn* o) {
  for (int l = 0; l < 3; ++l)
    hipLaunchKernelGGL(gd::nn_pack_kernel, dim3((27 * 256 + 255) / 256), dim3(256), 0, o->stream,
                       o->p + (size_t)l * gd::kLayerParams, o->wf + (size_t)l * 27 * 256,
                       o->wb + (size_t)l * 27 * 256, kCin[l], kCout[l]);
  NH(o, hipGetLastError());
  return GD_OK;
}

int forward_on(gd_nn* o, hipStream_t s) {
  const gd::NnGrid g = nn_grid(o);
  const unsigned nb = conv_blocks(o);
  const size_t L = gd::kLayerParams;
  hipLaunchKernelGGL((gd::conv3_lds_kernel<1, 0>), dim3(nb), dim3(256), 0, s, o->x, o->wf,
                     o->p + 27 * 256, (void*)o->a1, (const uint16_t*)nullptr, g);
  hipLaunchKernelGGL((gd::conv3_lds_kernel<1, 0>), dim3(nb), dim3(256), 0, s, o->a1,
                     o->wf + 27 * 256, o->p + L + 27 * 256, (void*)o->a2, (const uint16_t*)nullptr,
                     g);
  hipLaunchKernelGGL((gd::conv3_lds_kernel<0, 1>), dim3(nb), dim3(256), 0, s, o->a2,
                     o->wf + 2 * 27 * 256, o->p + 2 * L + 27 * 256, (void*)o->y,
                     (const uint16_t*)nullptr, g);
  NH(o, hipGetLastError());
  return GD_OK;
}

int forward(gd_nn* o) { return forward_on(o, o->stream); }

int wgrad(gd_nn* o, const uint16_t* dz, const uint16_t* act, float* gout) {
  const gd::NnGrid g = nn_grid(o);
  hipLaunchKernelGGL(gd::conv3_wgrad_kernel, dim3((unsigned)((o->nwg + 3) / 4)), dim3(256), 0,
                     o->stream, dz, act, o->part, g, o->ncell, o->chunk);
  hipLaunchKernelGGL(gd::nn_reduce_kernel, dim3((gd::kLayerParams + 255) / 256), dim3(256), 0,
                     o->stream, (const float*)o->part, o->nwg, gd::kLayerParams, gout);
  NH(o, hipGetLastError());
  return GD_OK;
}

}  // namespace

extern "C" {

int gd_nn_create(gd_nn** out, size_t nx, size_t ny, size_t nz) {
  if (!out) return GD_ERR_ARG;
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 4096 || nz > 4096) {
    g_nn_create_error = "grid dims must be in 1..4096";
    return GD_ERR_ARG;
  }
  gd_nn* o = new gd_nn();
  o->nx = (int)nx;
  o->ny = (int)ny;
  o->nz = (int)nz;
  o->ncell = (long long)nx * ny * nz;
  const size_t nc = (size_t)o->ncell;
  o->chunk = 1024;
  o->nwg = (int)((o->ncell + o->chunk - 1) / o->chunk);
  o->nloss = (int)((o->ncell + 255) / 256);
  bool ok = hipStreamCreate(&o->stream) == hipSuccess;
  auto alloc = [&](void** p, size_t bytes) {
    if (ok) ok = hipMalloc(p, bytes) == hipSuccess;
    if (ok) ok = hipMemset(*p, 0, bytes) == hipSuccess;
  };
  alloc((void**)&o->x, nc * 32);
  alloc((void**)&o->a1, nc * 32);
  alloc((void**)&o->a2, nc * 32);
  alloc((void**)&o->y, nc * 64);
  alloc((void**)&o->target, nc * 64);
  alloc((void**)&o->mask, nc);
  alloc((void**)&o->dz3, nc * 32);
  alloc((void**)&o->dz2, nc * 32);
  alloc((void**)&o->dz1, nc * 32);
  alloc((void**)&o->p, kNParams * 4);
  alloc((void**)&o->g, kNParams * 4);
  alloc((void**)&o->m1, kNParams * 4);
  alloc((void**)&o->m2, kNParams * 4);
  alloc((void**)&o->wf, 3 * 27 * 256 * 2);
  alloc((void**)&o->wb, 3 * 27 * 256 * 2);
  alloc((void**)&o->part, (size_t)o->nwg * gd::kLayerParams * 4);
  alloc((void**)&o->loss_part, (size_t)o->nloss * 4);
  if (!ok) {
    g_nn_create_error = "device allocation failed";
    gd_nn_destroy(o);
    return GD_ERR_NOMEM;
  }
  *out = o;
  return GD_OK;
}

int gd_nn_destroy(gd_nn* o) {
  if (!o) return GD_OK;
  if (o->stream) (void)hipStreamSynchronize(o->stream);
  void* ptrs[] = {o->x, o->a1, o->a2, o->y, o->target, o->mask, o->dz3, o->dz2, o->dz1,
                  o->p, o->g, o->m1, o->m2, o->wf, o->wb, o->part, o->loss_part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (o->stream) (void)hipStreamDestroy(o->stream);
  delete o;
  return GD_OK;
}

const char* gd_nn_last_error(const gd_nn* o) { return o ? o->err.c_str() : g_nn_create_error.c_str(); }

int gd_nn_num_params(size_t* n) {
  if (!n) return GD_ERR_ARG;
  *n = kNParams;
  return GD_OK;
}

int gd_nn_set_params(gd_nn* o, const float* params, size_t n) {
  if (!o) return GD_ERR_ARG;
  if (!params || n != (size_t)kNParams) return o->fail(GD_ERR_ARG, "expected 3*6928 parameters");
  NH(o, hipMemcpyAsync(o->p, params, n * 4, hipMemcpyHostToDevice, o->stream));
  NH(o, hipMemsetAsync(o->m1, 0, n * 4, o->stream));
  NH(o, hipMemsetAsync(o->m2, 0, n * 4, o->stream));
  o->adam_t = 0;
  int rc = pack_weights(o);
  if (rc != GD_OK) return rc;
  NH(o, hipStreamSynchronize(o->stream));
  return GD_OK;
}

int gd_nn_get_params(gd_nn* o, float* params, size_t n) {
  if (!o) return GD_ERR_ARG;
  if (!params || n != (size_t)kNParams) return o->fail(GD_ERR_ARG, "expected 3*6928 parameters");
  NH(o, hipStreamSynchronize(o->stream));
  NH(o, hipMemcpy(params, o->p, n * 4, hipMemcpyDeviceToHost));
  return GD_OK;
}

// Input features (fp32 [cell][16], rounded to bf16 here), training targets
// (fp32 [cell][16]) and the loss mask (1 byte per cell).
int gd_nn_set_input(gd_nn* o, const float* x) {
  if (!o || !x) return GD_ERR_ARG;
  std::vector<uint16_t> h((size_t)o->ncell * 16);
  for (size_t q = 0; q < h.size(); ++q) h[q] = gd::f32_to_bf16(x[q]);
  NH(o, hipMemcpyAsync(o->x, h.data(), h.size() * 2, hipMemcpyHostToDevice, o->stream));
  NH(o, hipStreamSynchronize(o->stream));
  return GD_OK;
}

int gd_nn_set_target(gd_nn* o, const float* t, const uint8_t* mask) {
  if (!o || !t || !mask) return GD_ERR_ARG;
  NH(o, hipMemcpyAsync(o->target, t, (size_t)o->ncell * 64, hipMemcpyHostToDevice, o->stream));
  NH(o, hipMemcpyAsync(o->mask, mask, (size_t)o->ncell, hipMemcpyHostToDevice, o->stream));
  NH(o, hipStreamSynchronize(o->stream));
  return GD_OK;
}

int gd_nn_forward(gd_nn* o, float* y_out) {
  if (!o) return GD_ERR_ARG;
  int rc = forward(o);
  if (rc != GD_OK) return rc;
  if (y_out) NH(o, hipMemcpyAsync(y_out, o->y, (size_t)o->ncell * 64, hipMemcpyDeviceToHost, o->stream));
  NH(o, hipStreamSynchronize(o->stream));
  return GD_OK;
}

// One training step on the current input/target: forward, MSE loss over the
// masked cells and the 12 output channels, backward, Adam. Returns the loss
// before the update. grads_out (optional, 3*6928 floats) receives the
// gradients.
int gd_nn_train_step(gd_nn* o, float lr, float* loss, float* grads_out) {
  if (!o) return GD_ERR_ARG;
  int rc = forward(o);
  if (rc != GD_OK) return rc;
  // the loss normalisation: number of masked cells x 12 channels
  std::vector<uint8_t> hm;
  long long count = 0;
  {
    hm.resize((size_t)o->ncell);
    NH(o, hipMemcpyAsync(hm.data(), o->mask, hm.size(), hipMemcpyDeviceToHost, o->stream));
    NH(o, hipStreamSynchronize(o->stream));
    for (auto v : hm) count += v != 0;
  }
  if (count == 0) return o->fail(GD_ERR_ARG, "empty loss mask");
  const float inv = 1.0f / (float)(count * 12);
  hipLaunchKernelGGL(gd::nn_loss_kernel, dim3((unsigned)o->nloss), dim3(256), 0, o->stream,
                     (const float*)o->y, (const float*)o->target, (const uint8_t*)o->mask,
                     o->ncell, 12, inv, o->dz3, o->loss_part);
  const gd::NnGrid g = nn_grid(o);
  const unsigned nb = conv_blocks(o);
  const size_t L = gd::kLayerParams;
  // layer 3
  rc = wgrad(o, o->dz3, o->a2, o->g + 2 * L);
  if (rc != GD_OK) return rc;
  hipLaunchKernelGGL((gd::conv3_lds_kernel<2, 0>), dim3(nb), dim3(256), 0, o->stream,
                     (const uint16_t*)o->dz3, (const uint16_t*)(o->wb + 2 * 27 * 256),
                     (const float*)nullptr, (void*)o->dz2, (const uint16_t*)o->a2, g);
  // layer 2
  rc = wgrad(o, o->dz2, o->a1, o->g + L);
  if (rc != GD_OK) return rc;
  hipLaunchKernelGGL((gd::conv3_lds_kernel<2, 0>), dim3(nb), dim3(256), 0, o->stream,
                     (const uint16_t*)o->dz2, (const uint16_t*)(o->wb + 27 * 256),
                     (const float*)nullptr, (void*)o->dz1, (const uint16_t*)o->a1, g);
  // layer 1
  rc = wgrad(o, o->dz1, o->x, o->g);
  if (rc != GD_OK) return rc;
  NH(o, hipGetLastError());
  std::vector<float> lp((size_t)o->nloss);
  NH(o, hipMemcpyAsync(lp.data(), o->loss_part, lp.size() * 4, hipMemcpyDeviceToHost, o->stream));
  if (grads_out) NH(o, hipMemcpyAsync(grads_out, o->g, kNParams * 4, hipMemcpyDeviceToHost, o->stream));
  // Adam
  o->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f / (1.0f - (float)std::pow(b1, (double)o->adam_t));
  const float c2 = 1.0f / (1.0f - (float)std::pow(b2, (double)o->adam_t));
  if (lr > 0.0f) {
    hipLaunchKernelGGL(gd::nn_adam_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, o->stream,
                       o->p, (const float*)o->g, o->m1, o->m2, kNParams, lr, b1, b2, 1e-8f, c1, c2);
    rc = pack_weights(o);
    if (rc != GD_OK) return rc;
  }
  NH(o, hipStreamSynchronize(o->stream));
  double s = 0.0;
  for (float v : lp) s += v;
  if (loss) *loss = (float)(s / (double)(count * 12));
  return GD_OK;
}

// Device check of the MFMA operand layout the convolutions assume: D = A B
// for 16x16 BF16 matrices given row-major as floats (rounded to bf16 here).
int gd_nn_mfma_selftest(const float* a, const float* b, float* d);

}  // extern "C"

namespace gd {
namespace {
__global__ void mfma_selftest_kernel(const uint16_t* a, const uint16_t* b, float* d) {
  const int l = threadIdx.x;
  bf16x4 av, bv;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    av[j] = (short)a[(l % 16) * 16 + 4 * (l / 16) + j];  // A[l%16][4*(l/16)+j]
    bv[j] = (short)b[(4 * (l / 16) + j) * 16 + (l % 16)];  // B[4*(l/16)+j][l%16]
  }
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = mfma16(av, bv, acc);
#pragma unroll
  for (int j = 0; j < 4; ++j) d[(4 * (l / 16) + j) * 16 + (l % 16)] = acc[j];
}
}  // namespace
}  // namespace gd

extern "C" int gd_nn_mfma_selftest(const float* a, const float* b, float* d) {
  if (!a || !b || !d) return GD_ERR_ARG;
  std::vector<uint16_t> ha(256), hb(256);
  for (int q = 0; q < 256; ++q) {
    ha[q] = gd::f32_to_bf16(a[q]);
    hb[q] = gd::f32_to_bf16(b[q]);
  }
  uint16_t *da = nullptr, *db = nullptr;
  float* dd = nullptr;
  if (hipMalloc(&da, 512) != hipSuccess || hipMalloc(&db, 512) != hipSuccess ||
      hipMalloc(&dd, 1024) != hipSuccess)
    return GD_ERR_NOMEM;
  hipError_t e = hipMemcpy(da, ha.data(), 512, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, hb.data(), 512, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(gd::mfma_selftest_kernel, dim3(1), dim3(64), 0, 0, da, db, dd);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(d, dd, 1024, hipMemcpyDeviceToHost);
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dd);
  return e == hipSuccess ? GD_OK : GD_ERR_HIP;
}

namespace gd {

hipError_t launch_restrict(const Fields& f, int nx, int ny, int nz, float* out, hipStream_t s) {
  const int cnx = nx / 2, cny = ny / 2, cnz = nz / 2;
  const long long cn = (long long)cnx * cny * cnz;
  if (cn == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(restrict_kernel, dim3((unsigned)((cn + 255) / 256)), dim3(256), 0, s, f, nx,
                     ny, nz, cnx, cny, cnz, out);
  return hipGetLastError();
}

hipError_t launch_expand_material(const uint8_t* id, const float* table, size_t n, float* out,
                                  hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(expand_material_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     id, table, (long long)n, out);
  return hipGetLastError();
}

int nn_dims(const gd_nn* o, int* nx, int* ny, int* nz) {
  *nx = o->nx;
  *ny = o->ny;
  *nz = o->nz;
  return GD_OK;
}

hipError_t nn_correct(gd_nn* o, const Fields& f, const float* pe0, const float* pe1,
                      const float* pe2, const float* cb, const NnFeat& nf, hipStream_t s) {
  const long long n = o->ncell;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(nn_features_kernel, dim3(nb), dim3(256), 0, s, f, pe0, pe1, pe2, cb, nf, o->nx,
                     o->ny, o->nz, o->x);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (forward_on(o, s) != GD_OK) return hipErrorLaunchFailure;
  hipLaunchKernelGGL(nn_apply_kernel, dim3(nb), dim3(256), 0, s, f, (const float*)o->y, nf, o->nx,
                     o->ny, o->nz);
  return hipGetLastError();
}

}  // namespace gd
