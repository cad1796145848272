// PLAIN stepping kernels for gfx950: one E launch, then one H launch per step,
// fields updated in place.
//
// Same physics and launch structure as the reference's naive path
// (updateE_3Dmap_fix / updateH_3Dmap_fix, kernels.cuh:337-394, two launches
// per step in update_FDTD_gpu_3D_warp_underutilization_fix,
// gdiamond_gpu.cuh:896-913). The mapping is MI355X-first instead of one thread
// per cell: a 64x4 workgroup covers a 64-wide x row segment (one wave64 per
// row, coalesced per array access) and each thread marches along z over a
// chunk of planes, reusing the z-neighbour it loaded on the previous plane
// from registers. The DFT monitor of E components is fused into the E kernel
// (the value after step n is final there); H components into the H kernel.
//
// Templated on the storage element S (float, double, or _Float16 / __bf16 with a
// per-component power-of-two scale; arithmetic runs in Codec<S>::T) and on the
// material kind (dense isotropic Cb in the compute type, the reference's 12
// arrays, or a 1-byte material id into a small FP32 Cb table that stays in
// cache).
#include "gd_common.hpp"

namespace gd {

namespace {

constexpr int kBx = 64;   // one wave64 along x
constexpr int kBy = 4;    // rows per workgroup
constexpr int kKc = 16;   // z planes per thread

template <class S, int kMat>
__global__ __launch_bounds__(kBx* kBy) void plain_e_kernel(Fields f, Scales sc, Material m, Grid g,
                                                           Sources s, Dft d, int n,
                                                           int run_step) {
  using C = Codec<S>;
  using T = typename C::T;
  const int i = blockIdx.x * kBx + threadIdx.x;
  const int j = blockIdx.y * kBy + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int k0 = blockIdx.z * kKc;
  const int k1 = min(g.nz, k0 + kKc);
  const bool ij_int = i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2;
  const bool dft_e = (d.mask & 7u) != 0u;
  if (!ij_int && !dft_e) return;

  S* __restrict__ Ex = static_cast<S*>(f.c[0]);
  S* __restrict__ Ey = static_cast<S*>(f.c[1]);
  S* __restrict__ Ez = static_cast<S*>(f.c[2]);
  const S* __restrict__ Hx = static_cast<const S*>(f.c[3]);
  const S* __restrict__ Hy = static_cast<const S*>(f.c[4]);
  const S* __restrict__ Hz = static_cast<const S*>(f.c[5]);
  const T* const* cf = reinterpret_cast<const T* const*>(m.c);
  const T* __restrict__ cbp = static_cast<const T*>(m.cb);
  const long long nx = g.nx, sxy = g.sxy;
  const T dx = (T)g.dx;
  long long idx = (long long)i + (long long)j * nx + (long long)k0 * sxy;
  auto hx_ = [&](long long q) { return C::dec(Hx[q], sc.s[3]); };
  auto hy_ = [&](long long q) { return C::dec(Hy[q], sc.s[4]); };
  auto hz_ = [&](long long q) { return C::dec(Hz[q], sc.s[5]); };

  // z-1 neighbours of the first plane; later ones come from the previous plane.
  T hx_km = 0, hy_km = 0;
  if (ij_int && k0 >= 1) {
    hx_km = hx_(idx - sxy);
    hy_km = hy_(idx - sxy);
  }
  for (int k = k0; k < k1; ++k, idx += sxy) {
    T ex = C::dec(Ex[idx], sc.s[0]), ey = C::dec(Ey[idx], sc.s[1]), ez = C::dec(Ez[idx], sc.s[2]);
    T hx = 0, hy = 0;
    if (ij_int) {
      hx = hx_(idx);
      hy = hy_(idx);
      if (k >= 1 && k <= g.nz - 2) {
        const T hz = hz_(idx);
        const T hz_jm = hz_(idx - nx);
        const T hz_im = hz_(idx - 1);
        const T hy_im = hy_(idx - 1);
        const T hx_jm = hx_(idx - nx);
        float jx, jy, jz;
        source_terms(s, 0, i, j, k, n, jx, jy, jz);
        const T cx = ((hz - hz_jm) - (hy - hy_km)) - (T)jx * dx;
        const T cy = ((hx - hx_km) - (hz - hz_im)) - (T)jy * dx;
        const T cz = ((hy - hy_im) - (hx - hx_jm)) - (T)jz * dx;
        if constexpr (kMat == kMatFull) {
          ex = cf[0][idx] * ex + cf[1][idx] * cx;
          ey = cf[2][idx] * ey + cf[3][idx] * cy;
          ez = cf[4][idx] * ez + cf[5][idx] * cz;
        } else {
          T b;
          if constexpr (kMat == kMatIdx)
            b = (T)m.table[m.id[idx]];
          else
            b = cbp[idx];
          ex = ex + b * cx;
          ey = ey + b * cy;
          ez = ez + b * cz;
        }
        Ex[idx] = C::enc(ex, sc.inv[0]);
        Ey[idx] = C::enc(ey, sc.inv[1]);
        Ez[idx] = C::enc(ez, sc.inv[2]);
        // the field value after the step is the stored one
        ex = C::round(ex, sc.s[0], sc.inv[0]);
        ey = C::round(ey, sc.s[1], sc.inv[1]);
        ez = C::round(ez, sc.s[2], sc.inv[2]);
      }
    }
    if (dft_e) {
      const long long p = dft_plane_index(d, g, i, j, k);
      if (p >= 0) dft_accumulate_group<true>(d, p, (double)ex, (double)ey, (double)ez, run_step);
    }
    hx_km = hx;
    hy_km = hy;
  }
}

template <class S, int kMat>
__global__ __launch_bounds__(kBx* kBy) void plain_h_kernel(Fields f, Scales sc, Material m, Grid g,
                                                           Sources s, Dft d, int n,
                                                           int run_step) {
  using C = Codec<S>;
  using T = typename C::T;
  const int i = blockIdx.x * kBx + threadIdx.x;
  const int j = blockIdx.y * kBy + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  const int k0 = blockIdx.z * kKc;
  const int k1 = min(g.nz, k0 + kKc);
  const bool ij_int = i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2;
  const bool dft_h = (d.mask & 56u) != 0u;
  if (!ij_int && !dft_h) return;

  const S* __restrict__ Ex = static_cast<const S*>(f.c[0]);
  const S* __restrict__ Ey = static_cast<const S*>(f.c[1]);
  const S* __restrict__ Ez = static_cast<const S*>(f.c[2]);
  S* __restrict__ Hx = static_cast<S*>(f.c[3]);
  S* __restrict__ Hy = static_cast<S*>(f.c[4]);
  S* __restrict__ Hz = static_cast<S*>(f.c[5]);
  const T* const* cf = reinterpret_cast<const T* const*>(m.c);
  const long long nx = g.nx, sxy = g.sxy;
  const T dx = (T)g.dx;
  const T db = sizeof(T) == 8 ? (T)m.db64 : (T)m.db;
  long long idx = (long long)i + (long long)j * nx + (long long)k0 * sxy;
  auto ex_ = [&](long long q) { return C::dec(Ex[q], sc.s[0]); };
  auto ey_ = [&](long long q) { return C::dec(Ey[q], sc.s[1]); };
  auto ez_ = [&](long long q) { return C::dec(Ez[q], sc.s[2]); };

  // own E of the current plane; the z+1 values loaded on a plane become the
  // own values of the next one.
  T ex = 0, ey = 0;
  if (ij_int) {
    ex = ex_(idx);
    ey = ey_(idx);
  }
  for (int k = k0; k < k1; ++k, idx += sxy) {
    T hx = C::dec(Hx[idx], sc.s[3]), hy = C::dec(Hy[idx], sc.s[4]), hz = C::dec(Hz[idx], sc.s[5]);
    T ex_kp = 0, ey_kp = 0;
    if (ij_int) {
      if (k + 1 < g.nz) {
        ex_kp = ex_(idx + sxy);
        ey_kp = ey_(idx + sxy);
      }
      if (k >= 1 && k <= g.nz - 2) {
        const T ez = ez_(idx);
        const T ez_jp = ez_(idx + nx);
        const T ez_ip = ez_(idx + 1);
        const T ex_jp = ex_(idx + nx);
        const T ey_ip = ey_(idx + 1);
        float mx, my, mz;
        source_terms(s, 3, i, j, k, n, mx, my, mz);
        const T cx = ((ey_kp - ey) - (ez_jp - ez)) - (T)mx * dx;
        const T cy = ((ez_ip - ez) - (ex_kp - ex)) - (T)my * dx;
        const T cz = ((ex_jp - ex) - (ey_ip - ey)) - (T)mz * dx;
        if constexpr (kMat == kMatFull) {
          hx = cf[6][idx] * hx + cf[7][idx] * cx;
          hy = cf[8][idx] * hy + cf[9][idx] * cy;
          hz = cf[10][idx] * hz + cf[11][idx] * cz;
        } else {
          hx = hx + db * cx;
          hy = hy + db * cy;
          hz = hz + db * cz;
        }
        Hx[idx] = C::enc(hx, sc.inv[3]);
        Hy[idx] = C::enc(hy, sc.inv[4]);
        Hz[idx] = C::enc(hz, sc.inv[5]);
        hx = C::round(hx, sc.s[3], sc.inv[3]);
        hy = C::round(hy, sc.s[4], sc.inv[4]);
        hz = C::round(hz, sc.s[5], sc.inv[5]);
      }
    }
    if (dft_h) {
      const long long p = dft_plane_index(d, g, i, j, k);
      if (p >= 0) dft_accumulate_group<false>(d, p, (double)hx, (double)hy, (double)hz, run_step);
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

template <class S, int kMat>
hipError_t launch_pair(const Fields& f, const Scales& sc, const Material& m, const Grid& g,
                       const Sources& s, const Dft& d, int n, int run_step, hipStream_t stream) {
  const dim3 block(kBx, kBy, 1);
  const dim3 grid((g.nx + kBx - 1) / kBx, (g.ny + kBy - 1) / kBy, (g.nz + kKc - 1) / kKc);
  hipLaunchKernelGGL((plain_e_kernel<S, kMat>), grid, block, 0, stream, f, sc, m, g, s, d, n,
                     run_step);
  hipLaunchKernelGGL((plain_h_kernel<S, kMat>), grid, block, 0, stream, f, sc, m, g, s, d, n,
                     run_step);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_plain_step(Storage st, const Fields& f, const Scales& sc, const Material& m,
                             MatKind mk, const Grid& g, const Sources& s, const Dft& d, int n,
                             int run_step, hipStream_t stream) {
  switch (st) {
    case kF32:
      if (mk == kMatFull) return launch_pair<float, kMatFull>(f, sc, m, g, s, d, n, run_step, stream);
      if (mk == kMatIdx) return launch_pair<float, kMatIdx>(f, sc, m, g, s, d, n, run_step, stream);
      return launch_pair<float, kMatIso>(f, sc, m, g, s, d, n, run_step, stream);
    case kF64:
      if (mk == kMatFull) return launch_pair<double, kMatFull>(f, sc, m, g, s, d, n, run_step, stream);
      if (mk == kMatIdx) return hipErrorInvalidValue;
      return launch_pair<double, kMatIso>(f, sc, m, g, s, d, n, run_step, stream);
    case kF16:
      if (mk == kMatFull) return hipErrorInvalidValue;
      if (mk == kMatIdx) return launch_pair<_Float16, kMatIdx>(f, sc, m, g, s, d, n, run_step, stream);
      return launch_pair<_Float16, kMatIso>(f, sc, m, g, s, d, n, run_step, stream);
    case kBF16:
      if (mk == kMatFull) return hipErrorInvalidValue;
      if (mk == kMatIdx) return launch_pair<__bf16, kMatIdx>(f, sc, m, g, s, d, n, run_step, stream);
      return launch_pair<__bf16, kMatIso>(f, sc, m, g, s, d, n, run_step, stream);
  }
  return hipErrorInvalidValue;
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
