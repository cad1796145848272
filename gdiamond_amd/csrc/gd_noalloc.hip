// gd_tiled_launch_noalloc: one TILED launch on caller-owned device buffers
// that allocates nothing.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops): the inputs and outputs are the caller's device arrays, and
// the launch's scratch memory -- the point-source descriptors, the DFT
// weights of the launch's steps and the per-level snapshot of the monitored
// plane -- lives in one file-scope device array, `scratch`, of fixed size.
// Each GPU holds its own copy of that array until the process ends; its
// address on the current device is looked up once per device
// (hipGetSymbolAddress) and only that address is ever handed to kernels.
//
// The launch runs the full TILED kernel (sources and DFT code in every tile)
// over every tile, so that no tile list is needed, followed by the snapshot
// fold of the DFT sums: the same kernels as gd_run's TILED launches, with the
// same results bit for bit.
//
// Calls on one GPU share that GPU's array: calls that overlap on one GPU (on
// different streams or threads) race on it. Nothing in it has to start at
// zero: every launch writes the descriptors and weights it reads, and every
// monitored plane cell's snapshot before the fold reads it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/gdiamond/gd_capi.h"
#include "gd_common.hpp"

namespace gd {
void dft_weight(double f_cps, long long n, double scale, double* wc, double* ws);
void set_call_error(const char* msg);
gd_tiling plan_tiling_dims(size_t nx, size_t ny, size_t nz, int storage);
}  // namespace gd

namespace {

constexpr size_t kParamBytes = 4096;
constexpr size_t kSnapFloats = (size_t)GD_NOALLOC_MAX_STEPS * GD_NOALLOC_MAX_SNAP_CELLS;
constexpr size_t kScratchBytes = kParamBytes + kSnapFloats * sizeof(float);

// The launch's scratch memory: [0, 4096) source descriptors and DFT weights
// (LaunchParams), then the DFT snapshot (levels x monitored components x
// plane floats).
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Source descriptors and DFT weights of one launch, passed by value to
// params_kernel, which writes them to the start of the scratch array.
struct LaunchParams {
  gd::SourceDesc src[GD_MAX_SOURCES];
  double wc[GD_NOALLOC_MAX_FREQS * GD_NOALLOC_MAX_STEPS];
  double ws[GD_NOALLOC_MAX_FREQS * GD_NOALLOC_MAX_STEPS];
};
static_assert(sizeof(LaunchParams) <= kParamBytes, "launch parameters exceed their slot");
constexpr size_t kOffWc = offsetof(LaunchParams, wc);
constexpr size_t kOffWs = offsetof(LaunchParams, ws);

__global__ __launch_bounds__(256) void params_kernel(LaunchParams p, unsigned char* dst) {
  const unsigned* s = reinterpret_cast<const unsigned*>(&p);
  unsigned* d = reinterpret_cast<unsigned*>(dst);
  for (unsigned q = threadIdx.x; q < sizeof(LaunchParams) / 4; q += blockDim.x) d[q] = s[q];
}

// Address of `scratch` on each device, looked up once per device with that
// device current.
std::mutex g_mu;
unsigned char* g_ptr[64] = {nullptr};

int scratch_ptr(unsigned char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    gd::set_call_error("no current HIP device");
    return GD_ERR_NODEVICE;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      gd::set_call_error((std::string("hipGetSymbolAddress(scratch): ") +
                          hipGetErrorString(e)).c_str());
      return GD_ERR_HIP;
    }
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return GD_OK;
}

int arg_error(const char* msg) {
  gd::set_call_error(msg);
  return GD_ERR_ARG;
}

}  // namespace

extern "C" {

int gd_tiled_launch_noalloc(const gd_noalloc_launch* L) {
  if (!L) return arg_error("null launch description");
  const int nx = L->nx, ny = L->ny, nz = L->nz;
  if (nx < 3 || ny < 3 || nz < 3) return arg_error("grid must be at least 3 cells on every axis");
  if ((unsigned long long)nx * (unsigned long long)ny * 4ull >= (1ull << 31))
    return arg_error("an xy plane must stay below 2 GiB");
  if (L->nsteps < 1 || L->nsteps > GD_NOALLOC_MAX_STEPS)
    return arg_error("nsteps must be 1..GD_NOALLOC_MAX_STEPS (one launch)");
  if (L->n0 < 0 || (long long)L->n0 + L->nsteps > (long long)INT32_MAX)
    return arg_error("step index out of range");
  for (int a = 0; a < 6; ++a) {
    if (!L->fields_in[a] || !L->fields_out[a]) return arg_error("null field array");
    for (int b = 0; b < 6; ++b)
      if (L->fields_out[a] == L->fields_in[b])
        return arg_error("fields_out must not be one of the fields_in arrays");
  }
  const bool indexed = L->material_id != nullptr;
  if (indexed ? !L->material_table : !L->cb) return arg_error("null material array");
  if (L->nsrc < 0 || L->nsrc > GD_MAX_SOURCES) return arg_error("0..GD_MAX_SOURCES sources");
  if (L->nsrc && !L->sources) return arg_error("null source array");
  for (int q = 0; q < L->nsrc; ++q) {
    const gd_point_source& s = L->sources[q];
    if (s.comp < 0 || s.comp > 5 || !s.table) return arg_error("bad source component or table");
    if (s.i < 0 || s.i >= nx || s.j < 0 || s.j >= ny || s.k < 0 || s.k >= nz)
      return arg_error("source outside grid");
  }
  int ncomp = 0;
  long long plane = 0;
  if (L->dft_mask) {
    if (L->dft_mask > 63u) return arg_error("DFT component mask must be in 0..63");
    if (L->dft_axis < 0 || L->dft_axis > 2) return arg_error("DFT axis must be 0, 1 or 2");
    const int lim = L->dft_axis == 0 ? nx : (L->dft_axis == 1 ? ny : nz);
    if (L->dft_index < 0 || L->dft_index >= lim) return arg_error("DFT plane outside grid");
    if (L->nf < 1 || L->nf > GD_NOALLOC_MAX_FREQS || !L->f_cycles_per_step)
      return arg_error("DFT needs 1..GD_NOALLOC_MAX_FREQS frequencies");
    if (!L->dft_re || !L->dft_im) return arg_error("null DFT sums");
    ncomp = __builtin_popcount(L->dft_mask);
    plane = L->dft_axis == 0 ? (long long)ny * nz
                             : (L->dft_axis == 1 ? (long long)nx * nz : (long long)nx * ny);
    if ((long long)ncomp * plane > GD_NOALLOC_MAX_SNAP_CELLS)
      return arg_error("monitored components x plane cells above GD_NOALLOC_MAX_SNAP_CELLS");
  }

  unsigned char* base = nullptr;
  const int rc = scratch_ptr(&base);
  if (rc != GD_OK) return rc;
  const hipStream_t stream = static_cast<hipStream_t>(L->stream);

  LaunchParams p;
  std::memset(&p, 0, sizeof(p));
  for (int q = 0; q < L->nsrc; ++q) {
    const gd_point_source& s = L->sources[q];
    p.src[q] = gd::SourceDesc{s.comp, s.i, s.j, s.k, s.table};
  }
  if (ncomp)
    for (int f = 0; f < L->nf; ++f)
      for (int r = 0; r < L->nsteps; ++r)
        gd::dft_weight(L->f_cycles_per_step[f], (long long)L->n0 + r, L->dft_scale,
                       &p.wc[f * L->nsteps + r], &p.ws[f * L->nsteps + r]);
  hipLaunchKernelGGL(params_kernel, dim3(1), dim3(256), 0, stream, p, base);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    gd::set_call_error((std::string("params_kernel: ") + hipGetErrorString(e)).c_str());
    return GD_ERR_HIP;
  }

  gd::Fields src, dst;
  for (int a = 0; a < 6; ++a) {
    src.c[a] = const_cast<float*>(L->fields_in[a]);  // read only
    dst.c[a] = L->fields_out[a];
  }
  gd::Scales sc;
  for (int a = 0; a < 6; ++a) sc.s[a] = sc.inv[a] = 1.0f;
  gd::Material m{};
  m.cb = L->cb;
  m.id = L->material_id;
  m.table = L->material_table;
  m.db = L->db;
  m.db64 = (double)L->db;
  gd::Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.sxy = (long long)nx * ny;
  g.dx = L->dx;
  gd::Sources s{};
  s.n = L->nsrc;
  s.d = reinterpret_cast<const gd::SourceDesc*>(base);
  gd::Dft d{};
  d.ncomp = ncomp;
  if (ncomp) {
    d.mask = L->dft_mask;
    d.axis = L->dft_axis;
    d.index = L->dft_index;
    d.nf = L->nf;
    d.run_n0 = L->n0;
    d.run_len = L->nsteps;
    d.plane = plane;
    d.wc = reinterpret_cast<const double*>(base + kOffWc);
    d.ws = reinterpret_cast<const double*>(base + kOffWs);
    d.re = L->dft_re;
    d.im = L->dft_im;
    d.snap = reinterpret_cast<float*>(base + kParamBytes);
  }
  const gd_tiling pt = gd::plan_tiling_dims((size_t)nx, (size_t)ny, (size_t)nz, GD_STORAGE_F32);
  const gd::ZTiling t{L->nsteps, pt.rows, pt.waves, pt.zchunk};
  e = gd::launch_zstream(t, gd::kF32, src, dst, sc, m, indexed ? gd::kMatIdx : gd::kMatIso, g, s,
                         d, L->n0, 0, gd::ZPart{gd::kZFull, nullptr, 0}, stream);
  if (e == hipSuccess && ncomp) e = gd::launch_dft_snapshot(d, L->nsteps, 0, stream);
  if (e != hipSuccess) {
    gd::set_call_error((std::string("TILED launch: ") + hipGetErrorString(e)).c_str());
    return GD_ERR_HIP;
  }
  return GD_OK;
}

}  // extern "C"
