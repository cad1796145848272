// PLAIN stepping kernels for gfx950: one E launch, then one H launch per step,
// fields updated in place.
//
// Same physics and launch structure as the reference's naive path
// (updateE_3Dmap_fix / updateH_3Dmap_fix, kernels.cuh:337-394, two launches
// per step in update_FDTD_gpu_3D_warp_underutilization_fix,
// gdiamond_gpu.cuh:896-913). The mapping is MI355X-first instead of one thread
// per cell: a 64x4 workgroup covers a 64-wide x row segment (one wave64 per
// row, 256 B coalesced per array access) and each thread marches along z over
// a chunk of planes, reusing the z-neighbour it loaded on the previous plane
// from registers. The DFT monitor of E components is fused into the E kernel
// (the value after step n is final there); H components into the H kernel.
#include "gd_common.hpp"

namespace gd {

namespace {

constexpr int kBx = 64;   // one wave64 along x
constexpr int kBy = 4;    // rows per workgroup
constexpr int kKc = 16;   // z planes per thread

template <bool kFull>
__global__ __launch_bounds__(kBx* kBy) void plain_e_kernel(Fields f, Material m, Grid g,
                                                           Sources s, Dft d, int n,
                                                           int run_step) {
  const int i = blockIdx.x * kBx + threadIdx.x;
  const int j = blockIdx.y * kBy + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int k0 = blockIdx.z * kKc;
  const int k1 = min(g.nz, k0 + kKc);
  const bool ij_int = i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2;
  const bool dft_e = (d.mask & 7u) != 0u;
  if (!ij_int && !dft_e) return;

  float* __restrict__ Ex = f.c[0];
  float* __restrict__ Ey = f.c[1];
  float* __restrict__ Ez = f.c[2];
  const float* __restrict__ Hx = f.c[3];
  const float* __restrict__ Hy = f.c[4];
  const float* __restrict__ Hz = f.c[5];
  const long long nx = g.nx, sxy = g.sxy;
  long long idx = (long long)i + (long long)j * nx + (long long)k0 * sxy;

  // z-1 neighbours of the first plane; later ones come from the previous plane.
  float hx_km = 0.0f, hy_km = 0.0f;
  if (ij_int && k0 >= 1) {
    hx_km = Hx[idx - sxy];
    hy_km = Hy[idx - sxy];
  }
  for (int k = k0; k < k1; ++k, idx += sxy) {
    float ex = Ex[idx], ey = Ey[idx], ez = Ez[idx];
    float hx = 0.0f, hy = 0.0f;
    if (ij_int) {
      hx = Hx[idx];
      hy = Hy[idx];
      if (k >= 1 && k <= g.nz - 2) {
        const float hz = Hz[idx];
        const float hz_jm = Hz[idx - nx];
        const float hz_im = Hz[idx - 1];
        const float hy_im = Hy[idx - 1];
        const float hx_jm = Hx[idx - nx];
        float jx, jy, jz;
        source_terms(s, 0, i, j, k, n, jx, jy, jz);
        const float cx = ((hz - hz_jm) - (hy - hy_km)) - jx * g.dx;
        const float cy = ((hx - hx_km) - (hz - hz_im)) - jy * g.dx;
        const float cz = ((hy - hy_im) - (hx - hx_jm)) - jz * g.dx;
        if (kFull) {
          ex = m.c[0][idx] * ex + m.c[1][idx] * cx;
          ey = m.c[2][idx] * ey + m.c[3][idx] * cy;
          ez = m.c[4][idx] * ez + m.c[5][idx] * cz;
        } else {
          const float b = m.cb[idx];
          ex = ex + b * cx;
          ey = ey + b * cy;
          ez = ez + b * cz;
        }
        Ex[idx] = ex;
        Ey[idx] = ey;
        Ez[idx] = ez;
      }
    }
    if (dft_e) {
      const long long p = dft_plane_index(d, g, i, j, k);
      if (p >= 0) dft_accumulate_group<true>(d, p, ex, ey, ez, run_step);
    }
    hx_km = hx;
    hy_km = hy;
  }
}

template <bool kFull>
__global__ __launch_bounds__(kBx* kBy) void plain_h_kernel(Fields f, Material m, Grid g,
                                                           Sources s, Dft d, int n,
                                                           int run_step) {
  const int i = blockIdx.x * kBx + threadIdx.x;
  const int j = blockIdx.y * kBy + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int k0 = blockIdx.z * kKc;
  const int k1 = min(g.nz, k0 + kKc);
  const bool ij_int = i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2;
  const bool dft_h = (d.mask & 56u) != 0u;
  if (!ij_int && !dft_h) return;

  const float* __restrict__ Ex = f.c[0];
  const float* __restrict__ Ey = f.c[1];
  const float* __restrict__ Ez = f.c[2];
  float* __restrict__ Hx = f.c[3];
  float* __restrict__ Hy = f.c[4];
  float* __restrict__ Hz = f.c[5];
  const long long nx = g.nx, sxy = g.sxy;
  long long idx = (long long)i + (long long)j * nx + (long long)k0 * sxy;

  // own E of the current plane; the z+1 values loaded on a plane become the
  // own values of the next one.
  float ex = 0.0f, ey = 0.0f;
  if (ij_int) {
    ex = Ex[idx];
    ey = Ey[idx];
  }
  for (int k = k0; k < k1; ++k, idx += sxy) {
    float hx = Hx[idx], hy = Hy[idx], hz = Hz[idx];
    float ex_kp = 0.0f, ey_kp = 0.0f;
    if (ij_int) {
      if (k + 1 < g.nz) {
        ex_kp = Ex[idx + sxy];
        ey_kp = Ey[idx + sxy];
      }
      if (k >= 1 && k <= g.nz - 2) {
        const float ez = Ez[idx];
        const float ez_jp = Ez[idx + nx];
        const float ez_ip = Ez[idx + 1];
        const float ex_jp = Ex[idx + nx];
        const float ey_ip = Ey[idx + 1];
        float mx, my, mz;
        source_terms(s, 3, i, j, k, n, mx, my, mz);
        const float cx = ((ey_kp - ey) - (ez_jp - ez)) - mx * g.dx;
        const float cy = ((ez_ip - ez) - (ex_kp - ex)) - my * g.dx;
        const float cz = ((ex_jp - ex) - (ey_ip - ey)) - mz * g.dx;
        if (kFull) {
          hx = m.c[6][idx] * hx + m.c[7][idx] * cx;
          hy = m.c[8][idx] * hy + m.c[9][idx] * cy;
          hz = m.c[10][idx] * hz + m.c[11][idx] * cz;
        } else {
          hx = hx + m.db * cx;
          hy = hy + m.db * cy;
          hz = hz + m.db * cz;
        }
        Hx[idx] = hx;
        Hy[idx] = hy;
        Hz[idx] = hz;
      }
    }
    if (dft_h) {
      const long long p = dft_plane_index(d, g, i, j, k);
      if (p >= 0) dft_accumulate_group<false>(d, p, hx, hy, hz, run_step);
    }
    ex = ex_kp;
    ey = ey_kp;
  }
}

__global__ __launch_bounds__(256) void copy_kernel(const float* __restrict__ src,
                                                   float* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x * 4;
  for (size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; i < n; i += stride) {
    if (i + 4 <= n) {
      *reinterpret_cast<float4*>(dst + i) = *reinterpret_cast<const float4*>(src + i);
    } else {
      for (size_t q = i; q < n; ++q) dst[q] = src[q];
    }
  }
}

}  // namespace

hipError_t launch_plain_step(const Fields& f, const Material& m, bool full, const Grid& g,
                             const Sources& s, const Dft& d, int n, int run_step,
                             hipStream_t stream) {
  const dim3 block(kBx, kBy, 1);
  const dim3 grid((g.nx + kBx - 1) / kBx, (g.ny + kBy - 1) / kBy, (g.nz + kKc - 1) / kKc);
  if (full) {
    hipLaunchKernelGGL(plain_e_kernel<true>, grid, block, 0, stream, f, m, g, s, d, n, run_step);
    hipLaunchKernelGGL(plain_h_kernel<true>, grid, block, 0, stream, f, m, g, s, d, n, run_step);
  } else {
    hipLaunchKernelGGL(plain_e_kernel<false>, grid, block, 0, stream, f, m, g, s, d, n, run_step);
    hipLaunchKernelGGL(plain_h_kernel<false>, grid, block, 0, stream, f, m, g, s, d, n, run_step);
  }
  return hipGetLastError();
}

hipError_t launch_copy(const float* src, float* dst, size_t n, hipStream_t stream) {
  const size_t per_block = 256 * 4;
  size_t blocks = (n + per_block - 1) / per_block;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, dst, n);
  return hipGetLastError();
}

}  // namespace gd
