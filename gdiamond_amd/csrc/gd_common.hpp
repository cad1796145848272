// Structures shared by the host dispatcher and the gfx950 stepping kernels.
//
// Arithmetic contract (all kernels): FP32, every operation rounded once (the
// library is compiled with -ffp-contract=off), operand order of the reference
// update (gdiamond.hpp:1193-1221):
//   E = Ca*E + Cb*(((Hp - Hq) - (Hr - Hs)) - J*dx)
//   H = Da*H + Db*(((Ep - Eq) - (Er - Es)) - M*dx)
// so results are bit-identical to the sequential CPU update.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gd {

constexpr int kMaxSources = 8;
constexpr int kMaxBlt = 8;

// Field arrays of one field copy (SoA, idx = i + j*nx + k*nx*ny).
struct Fields {
  float* c[6];  // Ex Ey Ez Hx Hy Hz
};

// Material. Full: 12 reference arrays Cax Cbx Cay Cby Caz Cbz Dax Dbx Day Dby
// Daz Dbz. Isotropic lossless: Ca = Da = 1 (multiplication by 1 is exact and
// omitted), cb per cell, db scalar.
struct Material {
  const float* c[12];
  const float* cb;
  float db;
};

// Point sources. The value table[n] enters the component's update at step n.
// The descriptors live in a small device array (not in the kernel arguments),
// so kernels index them at run time without holding them in registers.
struct SourceDesc {
  int comp, i, j, k;
  const float* table;  // device, indexed by absolute step
};
struct Sources {
  int n;
  const SourceDesc* d;  // device array of n descriptors
};

// DFT monitor state (FP64). Weights for run step r (absolute step run_n0 + r)
// are wc[f*run_len + r] = cos(2 pi f n)*scale and ws likewise with sin.
struct Dft {
  int ncomp;         // 0 = disabled
  unsigned mask;     // monitored components (bit c); slot of c = popcount(mask & ((1<<c)-1))
  int axis;          // 0: i = index, 1: j = index, 2: k = index
  int index;
  int nf;
  int run_n0;
  int run_len;
  long long plane;   // cells per monitored plane
  const double* wc;
  const double* ws;
  double* re;        // [slot][f][plane]
  double* im;
  float* snap;       // TILED: plane values per level, [level][component][plane]
};

struct Grid {
  int nx, ny, nz;
  long long sxy;     // nx*ny
  float dx;
};

// Plane index of (i,j,k) on the monitored plane, or -1 when off the plane.
__device__ __forceinline__ long long dft_plane_index(const Dft& d, const Grid& g, int i, int j,
                                                     int k) {
  if (d.axis == 2) return k == d.index ? (long long)i + (long long)j * g.nx : -1;
  if (d.axis == 0) return i == d.index ? (long long)j + (long long)k * g.ny : -1;
  return j == d.index ? (long long)i + (long long)k * g.nx : -1;
}

// re += v*wc; im -= v*ws for every frequency, for the component in `slot`.
__device__ __forceinline__ void dft_accumulate(const Dft& d, int slot, long long p, float v,
                                               int run_step) {
  const double dv = (double)v;
  for (int f = 0; f < d.nf; ++f) {
    const long long o = ((long long)slot * d.nf + f) * d.plane + p;
    const double wc = d.wc[(long long)f * d.run_len + run_step];
    const double ws = d.ws[(long long)f * d.run_len + run_step];
    d.re[o] = d.re[o] + dv * wc;
    d.im[o] = d.im[o] - dv * ws;
  }
}

// Accumulate the monitored components of one group (E: v = Ex,Ey,Ez; H:
// v = Hx,Hy,Hz) of one cell. Kept out of line: it runs only on the monitored
// plane, and inlining its scalar state would crowd the stepping loop.
template <bool kE>
__device__ __forceinline__ void dft_accumulate_group(const Dft& d, long long p, float v0,
                                                               float v1, float v2, int run_step) {
  const int base = kE ? 0 : 3;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if ((d.mask >> (base + c)) & 1u) {
      const int slot = __builtin_popcount(d.mask & ((1u << (base + c)) - 1u));
      dft_accumulate(d, slot, p, c == 0 ? v0 : (c == 1 ? v1 : v2), run_step);
    }
  }
}

// Source values of the E (J) or H (M) components of one cell at step n.
// Cells without a source get +0, and x - (+0)*dx == x exactly.
__device__ __forceinline__ void source_terms(const Sources& s, int base_comp, int i, int j, int k,
                                             int n, float& v0, float& v1, float& v2) {
  v0 = 0.0f;
  v1 = 0.0f;
  v2 = 0.0f;
  for (int q = 0; q < s.n; ++q) {
    const SourceDesc sd = s.d[q];
    if (sd.i == i && sd.j == j && sd.k == k) {
      const int c = sd.comp - base_comp;
      if (c >= 0 && c < 3) {
        const float v = sd.table[n];
        if (c == 0) v0 = v;
        if (c == 1) v1 = v;
        if (c == 2) v2 = v;
      }
    }
  }
}

}  // namespace gd

// Host-side launchers (defined in the .hip kernel files).
namespace gd {
hipError_t launch_plain_step(const Fields& f, const Material& m, bool full, const Grid& g,
                             const Sources& s, const Dft& d, int n, int run_step,
                             hipStream_t stream);

struct ZTiling {
  int blt, rows, waves, zchunk;
};
bool ztiling_supported(const ZTiling& t, bool full);
hipError_t launch_zstream(const ZTiling& t, const Fields& src, const Fields& dst, const Material& m,
                          bool full, const Grid& g, const Sources& s, const Dft& d, int n0,
                          int run_step0, hipStream_t stream);

hipError_t launch_copy(const float* src, float* dst, size_t n, hipStream_t stream);

// Accumulates the DFT of `nlev` consecutive steps from the per-level plane
// snapshot the TILED kernel wrote (run steps run_step0 .. run_step0+nlev-1).
hipError_t launch_dft_snapshot(const Dft& d, int nlev, int run_step0, hipStream_t stream);
}  // namespace gd
