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
struct Sources {
  int n;
  int comp[kMaxSources];
  int i[kMaxSources], j[kMaxSources], k[kMaxSources];
  const float* table[kMaxSources];  // device, indexed by absolute step
};

// DFT monitor state (FP64). Weights for run step r (absolute step run_n0 + r)
// are wc[f*run_len + r] = cos(2 pi f n)*scale and ws likewise with sin.
struct Dft {
  int ncomp;         // 0 = disabled
  int comp[6];       // monitored components, slot order
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

// Source values of the E (J) or H (M) components of one cell at step n.
// Cells without a source get +0, and x - (+0)*dx == x exactly.
__device__ __forceinline__ void source_terms(const Sources& s, int base_comp, int i, int j, int k,
                                             int n, float& v0, float& v1, float& v2) {
  v0 = 0.0f;
  v1 = 0.0f;
  v2 = 0.0f;
  for (int q = 0; q < s.n; ++q) {
    if (s.i[q] == i && s.j[q] == j && s.k[q] == k) {
      const int c = s.comp[q] - base_comp;
      if (c >= 0 && c < 3) {
        const float v = s.table[q][n];
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
}  // namespace gd
