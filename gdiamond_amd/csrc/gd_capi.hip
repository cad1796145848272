// C ABI of libgdiamond_hip.so: device-resident FDTD context, material,
// sources, DFT monitor and the stepping dispatcher.
//
// Replaces the per-call alloc/copy/run/free drivers of the reference
// (update_FDTD_gpu_2D gdiamond_gpu.cuh:483-648,
// update_FDTD_gpu_3D_warp_underutilization_fix :809-952,
// update_FDTD_gpu_figures :24-281, update_FDTD_mix_mapping_gpu_ver2/ver3
// gdiamond_gpu_mm_ver2.cuh:837, gdiamond_gpu_mm_ver3.cuh:478): device memory
// persists across calls, errors are status codes, and sources / DFT monitors
// are explicit instead of hard-wired.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gdiamond/gd_capi.h"
#include "gd_common.hpp"
#include "gd_half.hpp"

struct gd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t nx = 0, ny = 0, nz = 0, N = 0;
  float dx = 1.0f;
  void* f[2][6] = {{nullptr}};  // field copies, element type per storage
  void* fblock[2] = {nullptr, nullptr};  // one allocation per copy holding its six arrays
  int cur = 0;
  size_t step = 0;
  int storage = GD_STORAGE_F32;
  int scale_exp[6] = {0, 0, 0, 0, 0, 0};
  size_t esz = sizeof(float);  // bytes per stored field value

  bool have_mat = false;
  int mat_req = GD_MATERIAL_AUTO;  // requested isotropic format
  gd::MatKind mat_kind = gd::kMatIso;
  void* coeff[12] = {nullptr};  // full material, compute type (float / double)
  void* cb = nullptr;           // dense isotropic coefficient, compute type (float / double)
  uint8_t* mat_id = nullptr;    // indexed isotropic: id per cell
  float* mat_table = nullptr;   // indexed isotropic: Cb per id (GD_MAX_MATERIALS floats)
  int ntable = 0;
  float db = 0.0f;

  struct Src {
    int comp;
    size_t i, j, k;
    float* table;
    size_t len;
  };
  std::vector<Src> src;
  gd::SourceDesc* src_desc = nullptr;         // device copy of the descriptors
  std::vector<gd::SourceDesc> src_desc_host;

  int dft_ncomp = 0;
  int dft_comp[6] = {0};
  int dft_axis = 2;
  size_t dft_index = 0;
  size_t dft_plane = 0;
  std::vector<double> dft_f;
  double dft_scale = 1.0;
  double* dft_re = nullptr;
  double* dft_im = nullptr;
  double* dft_w = nullptr;
  size_t dft_w_cap = 0;
  float* dft_snap = nullptr;  // TILED: [level][slot][plane] plane values
  int dft_snap_levels = 0;    // levels dft_snap holds (set when allocated)
  std::vector<double> dft_w_host;

  // a run enqueued by gd_run_async that gd_sync (or the next call on the
  // context) has not completed yet: its launch-event count and length
  bool pending = false;
  size_t pending_nev = 0;
  size_t pending_nsteps = 0;
  int pending_variant = 0;

  // TILED launches split in two: tiles with sources or monitored cells run
  // the full kernel on `side`, the others the lean kernel on `stream`
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  struct Split {
    int blt = 0;  // tiling the classification is for (0: none yet in this run)
    std::vector<int> ids;  // the extra tiles' ids, then the lean items
    size_t nextra = 0;
    int* d_ids = nullptr;
    size_t cap = 0;  // ids the device buffer holds
    std::vector<gd::ZPack> packs;  // workgroups of remainder tiles (lean items -1-i)
    gd::ZPack* d_packs = nullptr;
    size_t pcap = 0;
  } split[2];

  bool launch_timing = false;
  std::vector<hipEvent_t> ev_pool;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_total_ms = 0.0, last_kernel_ms = 0.0;
  size_t last_nlaunch = 0;

  float* prev_e[3] = {nullptr, nullptr, nullptr};  // E one step back (learned corrector)
  float* cb_dense = nullptr;  // corrector: Cb per cell of an indexed material
  float* restrict_buf = nullptr;
  float* cal[2] = {nullptr, nullptr};
  size_t cal_n = 0;

  // pinned host staging for large device -> host field copies: two slots
  // used in turn, so one chunk's DMA overlaps the host copy-out of the last
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};

  std::string err;

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
  int hip_fail(hipError_t e, const char* what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return GD_ERR_HIP;
  }
};

#define GD_HIP(ctx, call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #call); \
  } while (0)

namespace {

thread_local std::string g_create_error;

// DFT weights use separate cos() and sin() calls that the compiler may not
// fuse into sincos(): glibc's sincos() differs from them in the last bit for
// some arguments, and the weights are part of the bit-exact contract.
__attribute__((noinline)) double weight_cos(double a) { return std::cos(a); }
__attribute__((noinline)) double weight_sin(double a) { return std::sin(a); }

// Completes a run enqueued by gd_run_async: waits for the context's stream
// and sets the run's timing (defined after gd_run). A no-op when none is
// pending. Every call that changes the context or reads its results calls it
// first, so the host buffers a run's uploads read are never refilled while
// the run may still read them.
int settle(gd_ctx* c);
#define GD_SETTLE(c)                    \
  do {                                  \
    const int settle_rc_ = settle(c);   \
    if (settle_rc_ != GD_OK) return settle_rc_; \
  } while (0)

// The six arrays of a field copy live in one allocation, each starting
// 16 KiB + 256 B after a 256-byte boundary past the previous one. With
// arrays a power-of-two size apart (512^3 FP32: 512 MiB), the same cell of
// the six components falls on the same memory channel and bank pattern;
// staggering them measured 222.0 -> 224.3 Gcell/s for 512^3 FP32 TILED and
// 280.8 -> 285.2 for binary16 (offsets of 0.25-256 KiB tried).
constexpr size_t kFieldStagger = 16640;

int alloc_fields(gd_ctx* c, int copy) {
  if (c->fblock[copy]) return GD_OK;
  const size_t bytes = c->N * c->esz;
  const size_t stride = (bytes + 255) / 256 * 256 + kFieldStagger;
  hipError_t e = hipMalloc(&c->fblock[copy], 6 * stride);
  if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(fields)");
  for (int a = 0; a < 6; ++a) {
    c->f[copy][a] = static_cast<char*>(c->fblock[copy]) + a * stride;
    GD_HIP(c, hipMemsetAsync(c->f[copy][a], 0, bytes, c->stream));
  }
  return GD_OK;
}

void free_fields(gd_ctx* c) {
  for (int k = 0; k < 2; ++k) {
    if (c->fblock[k]) (void)hipFree(c->fblock[k]);
    c->fblock[k] = nullptr;
    for (int a = 0; a < 6; ++a) c->f[k][a] = nullptr;
  }
}

float scale_of(const gd_ctx* c, int comp) { return std::ldexp(1.0f, c->scale_exp[comp]); }

// 16-bit storage formats (binary16, bfloat16) and their element conversions.
bool is16(int storage) { return storage == GD_STORAGE_F16 || storage == GD_STORAGE_BF16; }
uint16_t enc16(int storage, float v) {
  return storage == GD_STORAGE_BF16 ? gd::f32_to_bf16(v) : gd::f32_to_f16(v);
}
float dec16(int storage, uint16_t h) {
  return storage == GD_STORAGE_BF16 ? gd::bf16_to_f32(h) : gd::f16_to_f32(h);
}

// Runs f(begin, end) over [0, n) split across up to 16 host threads (one
// range when n is small). Used for the element-wise format conversions of
// whole fields, which are independent per element.
template <class F>
void parallel_for(size_t n, F&& f) {
  constexpr size_t kMinChunk = size_t(1) << 20;
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, n / kMinChunk));
  if (nt <= 1) {
    f(size_t(0), n);
    return;
  }
  const size_t chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b < e) th.emplace_back([&f, b, e] { f(b, e); });
  }
  f(size_t(0), std::min(n, chunk));
  for (auto& x : th) x.join();
}

// Device -> host copy of `bytes` in chunks of kStageBytes through the two
// pinned staging slots, handing each chunk to take(chunk, offset, len) on the
// host while the DMA of the next one runs. Pageable destinations copy at a
// fraction of the link rate (512^3 FP32: 9.5 GB/s); small copies go direct.
constexpr size_t kStageBytes = size_t(64) << 20;
constexpr size_t kStageMin = size_t(8) << 20;

int ensure_staging(gd_ctx* c) {
  for (int k = 0; k < 2; ++k) {
    if (!c->stage[k]) {
      hipError_t e = hipHostMalloc(&c->stage[k], kStageBytes, hipHostMallocDefault);
      if (e != hipSuccess) {
        c->stage[k] = nullptr;
        return c->hip_fail(e, "hipHostMalloc(staging)");
      }
    }
    if (!c->stage_ev[k]) {
      hipError_t e = hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming);
      if (e != hipSuccess) {
        c->stage_ev[k] = nullptr;
        return c->hip_fail(e, "hipEventCreate(staging)");
      }
    }
  }
  return GD_OK;
}

template <class Take>
int download(gd_ctx* c, const void* src, size_t bytes, Take&& take) {
  const int rc = ensure_staging(c);
  if (rc != GD_OK) return rc;
  const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
  const unsigned char* s = static_cast<const unsigned char*>(src);
  // chunk i goes to slot i & 1; slot (i + 1) & 1 is free again once the host
  // has taken chunk i - 1, which precedes the issue of chunk i + 1
  auto issue = [&](size_t i) -> hipError_t {
    const size_t off = i * kStageBytes, len = std::min(kStageBytes, bytes - off);
    hipError_t e =
        hipMemcpyAsync(c->stage[i & 1], s + off, len, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->stage_ev[i & 1], c->stream);
    return e;
  };
  if (nch) GD_HIP(c, issue(0));
  for (size_t i = 0; i < nch; ++i) {
    if (i + 1 < nch) GD_HIP(c, issue(i + 1));
    GD_HIP(c, hipEventSynchronize(c->stage_ev[i & 1]));
    const size_t off = i * kStageBytes, len = std::min(kStageBytes, bytes - off);
    take(static_cast<const unsigned char*>(c->stage[i & 1]), off, len);
  }
  return GD_OK;
}

// Host -> device copy of `bytes` from host memory: chunks of 8 MiB and more
// go through the two pinned staging slots (the host threads fill one while
// the DMA of the other runs), smaller copies go direct. Returns after the
// data is on the device.
int copy_in(gd_ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes < kStageMin) {
    GD_HIP(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GD_OK;
  }
  const int rc = ensure_staging(c);
  if (rc != GD_OK) return rc;
  const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
  const unsigned char* h = static_cast<const unsigned char*>(src);
  unsigned char* d = static_cast<unsigned char*>(dst);
  // on an error the copies already issued are drained before returning, so
  // that no DMA still reads a staging slot that a later transfer refills
  auto fail = [&](hipError_t e, const char* what) {
    (void)hipStreamSynchronize(c->stream);
    return c->hip_fail(e, what);
  };
  for (size_t i = 0; i < nch; ++i) {
    const int k = (int)(i & 1);
    // slot k was last read by the DMA of chunk i - 2
    if (i >= 2) {
      const hipError_t e = hipEventSynchronize(c->stage_ev[k]);
      if (e != hipSuccess) return fail(e, "hipEventSynchronize(staging)");
    }
    const size_t off = i * kStageBytes, len = std::min(kStageBytes, bytes - off);
    unsigned char* st = static_cast<unsigned char*>(c->stage[k]);
    parallel_for(len, [&](size_t b, size_t e) { std::memcpy(st + b, h + off + b, e - b); });
    hipError_t e = hipMemcpyAsync(d + off, st, len, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->stage_ev[k], c->stream);
    if (e != hipSuccess) return fail(e, "hipMemcpyAsync(staging)");
  }
  GD_HIP(c, hipStreamSynchronize(c->stream));
  return GD_OK;
}

// Host values -> storage bytes of one component (fp32 host values) for the
// FP64 and 16-bit formats (FP32 values are uploaded as they are).
void encode_component(const gd_ctx* c, int comp, const float* in, std::vector<unsigned char>& out) {
  out.resize(c->N * c->esz);
  if (c->storage == GD_STORAGE_F64) {
    double* o = reinterpret_cast<double*>(out.data());
    parallel_for(c->N, [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) o[q] = (double)in[q];
    });
  } else {
    const float inv = std::ldexp(1.0f, -c->scale_exp[comp]);
    const int st = c->storage;
    uint16_t* o = reinterpret_cast<uint16_t*>(out.data());
    parallel_for(c->N, [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) o[q] = enc16(st, in[q] * inv);
    });
  }
}

// Storage bytes -> double values (exact for every format).
void decode_to_f64(const gd_ctx* c, int comp, const unsigned char* in, double* out, size_t n) {
  if (c->storage == GD_STORAGE_F32) {
    const float* v = reinterpret_cast<const float*>(in);
    parallel_for(n, [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) out[q] = v[q];
    });
  } else if (c->storage == GD_STORAGE_F64) {
    std::memcpy(out, in, n * sizeof(double));
  } else {
    const float s = scale_of(c, comp);
    const int st = c->storage;
    const uint16_t* v = reinterpret_cast<const uint16_t*>(in);
    parallel_for(n, [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) out[q] = (double)(dec16(st, v[q]) * s);
    });
  }
}

gd::Scales scales_of(const gd_ctx* c) {
  gd::Scales sc;
  for (int a = 0; a < 6; ++a) {
    sc.s[a] = std::ldexp(1.0f, c->scale_exp[a]);
    sc.inv[a] = std::ldexp(1.0f, -c->scale_exp[a]);
  }
  return sc;
}

void free_material(gd_ctx* c) {
  for (int a = 0; a < 12; ++a) {
    if (c->coeff[a]) (void)hipFree(c->coeff[a]);
    c->coeff[a] = nullptr;
  }
  if (c->cb) (void)hipFree(c->cb);
  c->cb = nullptr;
  if (c->mat_id) (void)hipFree(c->mat_id);
  c->mat_id = nullptr;
  if (c->mat_table) (void)hipFree(c->mat_table);
  c->mat_table = nullptr;
  c->ntable = 0;
  c->have_mat = false;
}

int upload_indexed(gd_ctx* c, const uint8_t* id, const float* table, int ntable, float db) {
  free_material(c);
  hipError_t e = hipMalloc(&c->mat_id, c->N);
  if (e == hipSuccess) e = hipMalloc(&c->mat_table, GD_MAX_MATERIALS * sizeof(float));
  if (e != hipSuccess) {
    free_material(c);
    return c->hip_fail(e, "hipMalloc(material ids)");
  }
  float padded[GD_MAX_MATERIALS] = {0.0f};
  std::memcpy(padded, table, (size_t)ntable * sizeof(float));
  const int rc = copy_in(c, c->mat_id, id, c->N);
  if (rc != GD_OK) return rc;
  GD_HIP(c, hipMemcpy(c->mat_table, padded, sizeof(padded), hipMemcpyHostToDevice));
  c->ntable = ntable;
  c->db = db;
  c->mat_kind = gd::kMatIdx;
  c->have_mat = true;
  return GD_OK;
}

gd::Fields fields_of(gd_ctx* c, int copy) {
  gd::Fields F;
  for (int a = 0; a < 6; ++a) F.c[a] = c->f[copy][a];
  return F;
}

hipEvent_t pool_event(gd_ctx* c, size_t i) {
  while (c->ev_pool.size() <= i) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[i];
}

// Tiling planner: 4 rows x 8 waves (a 64 x 32 input tile), the fastest
// measured shape on MI355X, with B levels per launch. The z chunk is at most
// 32 owned planes for 4- and 8-byte fields and 64 for 16-bit fields, and
// shrinks (down to 16) until the launch has ~1024 workgroups (4 per CU on 256
// CUs).
gd_tiling plan_blt(size_t nx, size_t ny, size_t nz, int storage, int blt) {
  gd_tiling t;
  t.blt = blt;
  t.rows = 4;
  t.waves = 8;
  const size_t ox = 64 - 2 * t.blt, oy = (size_t)(t.rows * t.waves - 2 * t.blt);
  const size_t txy = ((nx + ox - 1) / ox) * ((ny + oy - 1) / oy);
  size_t zc = nz * txy / 1024;
  if (zc < 16) zc = 16;
  // FP32 with the lean kernel: 32-plane chunks run 0.5-1.7 % faster than 64
  // (512^3 220-222 vs 219-221, 768^3 225 vs 223), 24 or 16 slower (211, 201);
  // 16-bit fields: 64 (512^3 285 vs 278 at 32; 1024^3 297 vs 278 at 128)
  const size_t zmax = is16(storage) ? 64 : 32;
  if (zc > zmax) zc = zmax;
  if (zc > nz) zc = nz;
  t.zchunk = (int)zc;
  return t;
}

gd_tiling plan_dims(size_t nx, size_t ny, size_t nz, int storage) {
  // Four levels per launch for 4- and 8-byte fields. With 16-bit fields the
  // HBM share of a launch is half as large and three levels run faster
  // (227 vs 185 Gcell/s at 512^3: less overlap, no register spills).
  return plan_blt(nx, ny, nz, storage, is16(storage) ? 3 : 4);
}

gd_tiling plan(const gd_ctx* c) { return plan_dims(c->nx, c->ny, c->nz, c->storage); }

}  // namespace

namespace gd {
// DFT weights of absolute step n: cos(2 pi f n)*scale and sin(2 pi f n)*scale
// in double, as every run evaluates them (shared with the allocation-free
// entry point so that both give the same sums).
void dft_weight(double f_cps, long long n, double scale, double* wc, double* ws) {
  const double ang = 2.0 * 3.14159265358979323846 * f_cps * (double)n;
  *wc = weight_cos(ang) * scale;
  *ws = weight_sin(ang) * scale;
}
// Error text of calls made without a context (gd_last_error(NULL)).
void set_call_error(const char* msg) { g_create_error = msg; }
// The planner's tiling (shared with the allocation-free entry point).
gd_tiling plan_tiling_dims(size_t nx, size_t ny, size_t nz, int storage) {
  return plan_dims(nx, ny, nz, storage);
}
}  // namespace gd

extern "C" {

const char* gd_version(void) { return "gdiamond-hip 0.1 gfx950"; }

int gd_device_count(int* count) {
  if (!count) return GD_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return GD_OK;
}

int gd_create(gd_ctx** out, size_t nx, size_t ny, size_t nz, float dx) {
  if (!out) return GD_ERR_ARG;
  *out = nullptr;
  if (nx < 3 || ny < 3 || nz < 3) {
    g_create_error = "grid must be at least 3 cells on every axis";
    return GD_ERR_ARG;
  }
  // at most 2^40 cells in all, so that cell counts and byte sizes cannot wrap
  if (nx > (1u << 30) || ny > (1u << 30) || nz > (1u << 30) || nx * ny > (1ull << 40) ||
      nz > (1ull << 40) / (nx * ny)) {
    g_create_error = "grid too large";
    return GD_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_create_error = "no HIP device visible";
    return GD_ERR_NODEVICE;
  }
  gd_ctx* c = new gd_ctx();
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->N = nx * ny * nz;
  c->dx = dx;
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    g_create_error = "HIP stream/event creation failed";
    gd_destroy(c);
    return GD_ERR_HIP;
  }
  int rc = alloc_fields(c, 0);
  if (rc == GD_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = GD_ERR_HIP;
  if (rc != GD_OK) {
    g_create_error = c->err;
    gd_destroy(c);
    return rc == GD_ERR_HIP ? GD_ERR_NOMEM : rc;
  }
  *out = c;
  return GD_OK;
}

int gd_destroy(gd_ctx* c) {
  if (!c) return GD_OK;
  if (c->pending) (void)settle(c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_fields(c);
  for (auto& p : c->prev_e)
    if (p) (void)hipFree(p);
  if (c->cb_dense) (void)hipFree(c->cb_dense);
  if (c->restrict_buf) (void)hipFree(c->restrict_buf);
  free_material(c);
  for (auto& s : c->src) (void)hipFree(s.table);
  if (c->src_desc) (void)hipFree(c->src_desc);
  if (c->dft_re) (void)hipFree(c->dft_re);
  if (c->dft_im) (void)hipFree(c->dft_im);
  if (c->dft_w) (void)hipFree(c->dft_w);
  if (c->dft_snap) (void)hipFree(c->dft_snap);
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (int k = 0; k < 2; ++k)
    if (c->cal[k]) (void)hipFree(c->cal[k]);
  for (auto& sp : c->split) {
    if (sp.d_ids) (void)hipFree(sp.d_ids);
    if (sp.d_packs) (void)hipFree(sp.d_packs);
  }
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->join_ev) (void)hipEventDestroy(c->join_ev);
  for (int k = 0; k < 2; ++k) {
    if (c->stage[k]) (void)hipHostFree(c->stage[k]);
    if (c->stage_ev[k]) (void)hipEventDestroy(c->stage_ev[k]);
  }
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GD_OK;
}

const char* gd_last_error(const gd_ctx* c) {
  if (!c) return g_create_error.c_str();
  return c->err.c_str();
}

// ---- storage ---------------------------------------------------------------

int gd_set_storage(gd_ctx* c, int storage, const int* scale_exp) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (storage != GD_STORAGE_F32 && storage != GD_STORAGE_F64 && !is16(storage))
    return c->fail(GD_ERR_ARG, "unknown storage format");
  if (c->have_mat) return c->fail(GD_ERR_STATE, "set the storage format before the material");
  int se[6] = {0, 0, 0, 0, 0, 0};
  if (scale_exp) {
    for (int a = 0; a < 6; ++a) {
      if (scale_exp[a] < -100 || scale_exp[a] > 100)
        return c->fail(GD_ERR_ARG, "scale exponents must be in -100..100");
      se[a] = scale_exp[a];
    }
  }
  GD_HIP(c, hipStreamSynchronize(c->stream));
  free_fields(c);
  c->storage = storage;
  std::memcpy(c->scale_exp, se, sizeof(se));
  c->esz = storage == GD_STORAGE_F64 ? 8 : (is16(storage) ? 2 : 4);
  c->cur = 0;
  c->step = 0;
  int rc = alloc_fields(c, 0);
  if (rc != GD_OK) return rc;
  GD_HIP(c, hipStreamSynchronize(c->stream));
  return GD_OK;
}

int gd_get_storage(const gd_ctx* c, int* storage, int* scale_exp) {
  if (!c || !storage) return GD_ERR_ARG;
  *storage = c->storage;
  if (scale_exp) std::memcpy(scale_exp, c->scale_exp, sizeof(c->scale_exp));
  return GD_OK;
}

// ---- material -------------------------------------------------------------

int gd_set_coefficients(gd_ctx* c, const float* const coeff[12]) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!coeff) return c->fail(GD_ERR_ARG, "null coefficient table");
  for (int a = 0; a < 12; ++a)
    if (!coeff[a]) return c->fail(GD_ERR_ARG, "null coefficient array");
  if (is16(c->storage))
    return c->fail(GD_ERR_UNSUPPORTED, "16-bit storage supports the isotropic material only");
  free_material(c);
  const size_t es = c->storage == GD_STORAGE_F64 ? sizeof(double) : sizeof(float);
  std::vector<double> tmp;
  for (int a = 0; a < 12; ++a) {
    hipError_t e = hipMalloc(&c->coeff[a], c->N * es);
    if (e != hipSuccess) {
      free_material(c);
      return c->hip_fail(e, "hipMalloc(coefficients)");
    }
    if (c->storage == GD_STORAGE_F64) {
      tmp.assign(coeff[a], coeff[a] + c->N);
      const int rc = copy_in(c, c->coeff[a], tmp.data(), c->N * es);
      if (rc != GD_OK) return rc;
    } else {
      const int rc = copy_in(c, c->coeff[a], coeff[a], c->N * es);
      if (rc != GD_OK) return rc;
    }
  }
  c->mat_kind = gd::kMatFull;
  c->have_mat = true;
  return GD_OK;
}

int gd_set_material_format(gd_ctx* c, int format) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (format != GD_MATERIAL_AUTO && format != GD_MATERIAL_DENSE && format != GD_MATERIAL_INDEXED)
    return c->fail(GD_ERR_ARG, "unknown material format");
  if (c->have_mat) return c->fail(GD_ERR_STATE, "set the material format before the material");
  c->mat_req = format;
  return GD_OK;
}

int gd_get_material_format(const gd_ctx* c, int* format, int* ntable) {
  if (!c || !format) return GD_ERR_ARG;
  if (!c->have_mat)
    *format = -1;
  else
    *format = c->mat_kind == gd::kMatFull ? GD_MATERIAL_FULL
              : c->mat_kind == gd::kMatIdx ? GD_MATERIAL_INDEXED
                                           : GD_MATERIAL_DENSE;
  if (ntable) *ntable = c->have_mat && c->mat_kind == gd::kMatIdx ? c->ntable : 0;
  return GD_OK;
}

int gd_set_isotropic(gd_ctx* c, const float* cb, float db) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!cb) return c->fail(GD_ERR_ARG, "null cb array");
  // AUTO: the indexed form when Cb has at most GD_MAX_MATERIALS distinct
  // values (storage other than FP64), else the dense one. Both give the same
  // results; the indexed one reads 1 byte of material per cell instead of 4
  // and, with the lean TILED kernel, runs faster (512^3 FP32 220 against 211
  // Gcell/s, 1024^3 226 against 210; FP16 1024^3 304 against 296).
  const bool want_index =
      c->mat_req == GD_MATERIAL_INDEXED ||
      (c->mat_req == GD_MATERIAL_AUTO && c->storage != GD_STORAGE_F64);
  if (want_index) {
    if (c->storage == GD_STORAGE_F64)
      return c->fail(GD_ERR_UNSUPPORTED, "the indexed material needs FP32 or 16-bit storage");
    gd::HostVec<uint8_t> id;
    std::vector<float> table;
    if (gd::index_materials(cb, c->N, id, table))
      return upload_indexed(c, id.data(), table.data(), (int)table.size(), db);
    if (c->mat_req == GD_MATERIAL_INDEXED)
      return c->fail(GD_ERR_UNSUPPORTED, "more than 256 distinct Cb values for the indexed "
                                         "material");
  }
  free_material(c);
  // the coefficient is kept in the compute type: FP32 for FP32 and 16-bit
  // fields (exact), FP64 for FP64 fields
  const size_t ces = c->storage == GD_STORAGE_F64 ? sizeof(double) : sizeof(float);
  hipError_t e = hipMalloc(&c->cb, c->N * ces);
  if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(cb)");
  if (c->storage == GD_STORAGE_F64) {
    std::vector<double> buf(c->N);
    parallel_for(c->N, [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) buf[q] = (double)cb[q];
    });
    const int rc = copy_in(c, c->cb, buf.data(), c->N * ces);
    if (rc != GD_OK) return rc;
  } else {
    const int rc = copy_in(c, c->cb, cb, c->N * ces);
    if (rc != GD_OK) return rc;
  }
  c->db = db;
  c->mat_kind = gd::kMatIso;
  c->have_mat = true;
  return GD_OK;
}

int gd_set_isotropic_indexed(gd_ctx* c, const uint8_t* id, const float* table, int ntable,
                             float db) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!id || !table) return c->fail(GD_ERR_ARG, "null id or table");
  if (ntable < 1 || ntable > GD_MAX_MATERIALS)
    return c->fail(GD_ERR_ARG, "ntable must be in 1..256");
  for (size_t q = 0; q < c->N; ++q)
    if ((int)id[q] >= ntable) return c->fail(GD_ERR_ARG, "material id out of range");
  if (c->storage == GD_STORAGE_F64 || c->mat_req == GD_MATERIAL_DENSE) {
    if (c->storage == GD_STORAGE_F64 && c->mat_req == GD_MATERIAL_INDEXED)
      return c->fail(GD_ERR_UNSUPPORTED, "the indexed material needs FP32 or 16-bit storage");
    std::vector<float> cb(c->N);
    for (size_t q = 0; q < c->N; ++q) cb[q] = table[id[q]];
    const int req = c->mat_req;
    c->mat_req = GD_MATERIAL_DENSE;
    const int rc = gd_set_isotropic(c, cb.data(), db);
    c->mat_req = req;
    return rc;
  }
  return upload_indexed(c, id, table, ntable, db);
}

int gd_set_isotropic_eps(gd_ctx* c, const float* eps_r, float dt) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!eps_r) return c->fail(GD_ERR_ARG, "null eps array");
  gd::HostVec<float> cb(c->N);
  const double ddt = dt, ddx = c->dx;
  std::atomic<bool> bad{false};
  parallel_for(c->N, [&](size_t b, size_t e) {
    for (size_t q = b; q < e; ++q) {
      if (!(eps_r[q] > 0.0f)) {
        bad.store(true, std::memory_order_relaxed);
        return;
      }
      cb[q] = (float)(ddt / ((double)eps_r[q] * ddx));
    }
  });
  if (bad.load()) return c->fail(GD_ERR_ARG, "eps_r must be positive");
  return gd_set_isotropic(c, cb.data(), (float)(ddt / ddx));
}

// ---- sources ----------------------------------------------------------------

int gd_add_source(gd_ctx* c, int comp, size_t i, size_t j, size_t k, const float* table,
                  size_t nsteps) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (comp < 0 || comp > 5) return c->fail(GD_ERR_ARG, "component must be 0..5");
  if (i >= c->nx || j >= c->ny || k >= c->nz) return c->fail(GD_ERR_ARG, "source outside grid");
  if (!table || nsteps == 0) return c->fail(GD_ERR_ARG, "empty source table");
  if (nsteps > (size_t)INT32_MAX) return c->fail(GD_ERR_ARG, "source table too long");
  if ((int)c->src.size() >= GD_MAX_SOURCES) return c->fail(GD_ERR_ARG, "too many sources");
  // One value per cell and component, as the reference's J/M arrays hold.
  for (const auto& o : c->src)
    if (o.comp == comp && o.i == i && o.j == j && o.k == k)
      return c->fail(GD_ERR_ARG, "a source on this cell and component already exists");
  gd_ctx::Src s{comp, i, j, k, nullptr, nsteps};
  hipError_t e = hipMalloc(&s.table, nsteps * sizeof(float));
  if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(source)");
  e = hipMemcpy(s.table, table, nsteps * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(s.table);
    return c->hip_fail(e, "hipMemcpy(source)");
  }
  c->src.push_back(s);
  return GD_OK;
}

int gd_clear_sources(gd_ctx* c) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  GD_HIP(c, hipStreamSynchronize(c->stream));
  for (auto& s : c->src) (void)hipFree(s.table);
  c->src.clear();
  return GD_OK;
}

// ---- fields -----------------------------------------------------------------

int gd_set_fields(gd_ctx* c, const float* const fields[6]) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!fields) return c->fail(GD_ERR_ARG, "null field table");
  for (int a = 0; a < 6; ++a)
    if (!fields[a]) return c->fail(GD_ERR_ARG, "null field array");
  GD_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<unsigned char> buf;
  for (int a = 0; a < 6; ++a) {
    // FP32 fields are the storage bytes themselves: no host-side conversion
    const void* src = fields[a];
    if (c->storage != GD_STORAGE_F32) {
      encode_component(c, a, fields[a], buf);
      src = buf.data();
    }
    const int rc = copy_in(c, c->f[c->cur][a], src, c->N * c->esz);
    if (rc != GD_OK) return rc;
  }
  return GD_OK;
}

int gd_get_field_f64(gd_ctx* c, int comp, double* out) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (comp < 0 || comp > 5 || !out) return c->fail(GD_ERR_ARG, "bad component or null output");
  GD_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<unsigned char> buf(c->N * c->esz);
  GD_HIP(c, hipMemcpy(buf.data(), c->f[c->cur][comp], c->N * c->esz, hipMemcpyDeviceToHost));
  decode_to_f64(c, comp, buf.data(), out, c->N);
  return GD_OK;
}

int gd_get_fields(gd_ctx* c, float* const fields[6]) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!fields) return c->fail(GD_ERR_ARG, "null field table");
  for (int a = 0; a < 6; ++a) {
    int rc = gd_get_field(c, a, fields[a]);
    if (rc != GD_OK) return rc;
  }
  return GD_OK;
}

int gd_get_field(gd_ctx* c, int comp, float* out) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (comp < 0 || comp > 5 || !out) return c->fail(GD_ERR_ARG, "bad component or null output");
  GD_HIP(c, hipStreamSynchronize(c->stream));
  const size_t bytes = c->N * c->esz;
  if (bytes >= kStageMin) {
    const int st = c->storage;
    const float sc = scale_of(c, comp);
    return download(c, c->f[c->cur][comp], bytes,
                    [&](const unsigned char* chunk, size_t off, size_t len) {
      const size_t q0 = off / c->esz, n = len / c->esz;
      if (st == GD_STORAGE_F32) {
        unsigned char* dst = reinterpret_cast<unsigned char*>(out) + off;
        parallel_for(len, [&](size_t b, size_t e) { std::memcpy(dst + b, chunk + b, e - b); });
      } else if (is16(st)) {
        // a 16-bit value times a power of two is exact in float
        const uint16_t* v = reinterpret_cast<const uint16_t*>(chunk);
        parallel_for(n, [&](size_t b, size_t e) {
          for (size_t q = b; q < e; ++q) out[q0 + q] = dec16(st, v[q]) * sc;
        });
      } else {
        const double* v = reinterpret_cast<const double*>(chunk);
        parallel_for(n, [&](size_t b, size_t e) {
          for (size_t q = b; q < e; ++q) out[q0 + q] = (float)v[q];
        });
      }
    });
  }
  if (c->storage == GD_STORAGE_F32) {
    GD_HIP(c, hipMemcpy(out, c->f[c->cur][comp], c->N * sizeof(float), hipMemcpyDeviceToHost));
    return GD_OK;
  }
  std::vector<unsigned char> buf(c->N * c->esz);
  GD_HIP(c, hipMemcpy(buf.data(), c->f[c->cur][comp], c->N * c->esz, hipMemcpyDeviceToHost));
  if (is16(c->storage)) {
    // a 16-bit value times a power of two is exact in float
    const float sc = scale_of(c, comp);
    const int st = c->storage;
    const uint16_t* v = reinterpret_cast<const uint16_t*>(buf.data());
    parallel_for(c->N, [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) out[q] = dec16(st, v[q]) * sc;
    });
  } else {
    const double* v = reinterpret_cast<const double*>(buf.data());
    parallel_for(c->N, [&](size_t b, size_t e) {
      for (size_t q = b; q < e; ++q) out[q] = (float)v[q];
    });
  }
  return GD_OK;
}

int gd_get_fields_f64(gd_ctx* c, double* const fields[6]) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!fields) return c->fail(GD_ERR_ARG, "null field table");
  for (int a = 0; a < 6; ++a) {
    int rc = gd_get_field_f64(c, a, fields[a]);
    if (rc != GD_OK) return rc;
  }
  return GD_OK;
}

int gd_get_plane_xy(gd_ctx* c, int comp, size_t k, float* out) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (comp < 0 || comp > 5 || !out || k >= c->nz)
    return c->fail(GD_ERR_ARG, "bad component, plane or null output");
  GD_HIP(c, hipStreamSynchronize(c->stream));
  const size_t P = c->nx * c->ny;
  std::vector<unsigned char> buf(P * c->esz);
  GD_HIP(c, hipMemcpy(buf.data(), static_cast<unsigned char*>(c->f[c->cur][comp]) + k * P * c->esz,
                      P * c->esz, hipMemcpyDeviceToHost));
  std::vector<double> v(P);
  decode_to_f64(c, comp, buf.data(), v.data(), P);
  for (size_t q = 0; q < P; ++q) out[q] = (float)v[q];
  return GD_OK;
}

int gd_reset(gd_ctx* c) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  for (int a = 0; a < 6; ++a)
    GD_HIP(c, hipMemsetAsync(c->f[c->cur][a], 0, c->N * c->esz, c->stream));
  const size_t dn = (size_t)c->dft_ncomp * c->dft_f.size() * c->dft_plane;
  if (dn) {
    GD_HIP(c, hipMemsetAsync(c->dft_re, 0, dn * sizeof(double), c->stream));
    GD_HIP(c, hipMemsetAsync(c->dft_im, 0, dn * sizeof(double), c->stream));
  }
  GD_HIP(c, hipStreamSynchronize(c->stream));
  c->step = 0;
  return GD_OK;
}

int gd_get_step(const gd_ctx* c, size_t* step) {
  if (!c || !step) return GD_ERR_ARG;
  *step = c->step;
  return GD_OK;
}

// ---- DFT --------------------------------------------------------------------

int gd_dft_disable(gd_ctx* c) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  GD_HIP(c, hipStreamSynchronize(c->stream));
  if (c->dft_re) (void)hipFree(c->dft_re);
  if (c->dft_im) (void)hipFree(c->dft_im);
  if (c->dft_snap) (void)hipFree(c->dft_snap);
  c->dft_re = c->dft_im = nullptr;
  c->dft_snap = nullptr;
  c->dft_snap_levels = 0;
  c->dft_ncomp = 0;
  c->dft_f.clear();
  c->dft_plane = 0;
  return GD_OK;
}

int gd_dft_config(gd_ctx* c, unsigned mask, int axis, size_t index, const double* f, size_t nf,
                  double scale) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (mask == 0 || mask > 63u) return c->fail(GD_ERR_ARG, "component mask must be in 1..63");
  if (axis < 0 || axis > 2) return c->fail(GD_ERR_ARG, "axis must be 0, 1 or 2");
  const size_t lim = axis == 0 ? c->nx : (axis == 1 ? c->ny : c->nz);
  if (index >= lim) return c->fail(GD_ERR_ARG, "plane index outside grid");
  if (!f || nf == 0 || nf > 4096) return c->fail(GD_ERR_ARG, "need 1..4096 frequencies");
  int rc = gd_dft_disable(c);
  if (rc != GD_OK) return rc;
  c->dft_ncomp = 0;
  for (int a = 0; a < 6; ++a)
    if (mask & (1u << a)) c->dft_comp[c->dft_ncomp++] = a;
  c->dft_axis = axis;
  c->dft_index = index;
  c->dft_plane = axis == 0 ? c->ny * c->nz : (axis == 1 ? c->nx * c->nz : c->nx * c->ny);
  c->dft_f.assign(f, f + nf);
  c->dft_scale = scale;
  const size_t dn = (size_t)c->dft_ncomp * nf * c->dft_plane;
  hipError_t e = hipMalloc(&c->dft_re, dn * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->dft_im, dn * sizeof(double));
  if (e != hipSuccess) {
    c->dft_ncomp = 0;
    return c->hip_fail(e, "hipMalloc(dft)");
  }
  GD_HIP(c, hipMemsetAsync(c->dft_re, 0, dn * sizeof(double), c->stream));
  GD_HIP(c, hipMemsetAsync(c->dft_im, 0, dn * sizeof(double), c->stream));
  GD_HIP(c, hipStreamSynchronize(c->stream));
  return GD_OK;
}

int gd_dft_get(gd_ctx* c, int comp, double* re, double* im) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!re || !im) return c->fail(GD_ERR_ARG, "null output");
  int slot = -1;
  for (int q = 0; q < c->dft_ncomp; ++q)
    if (c->dft_comp[q] == comp) slot = q;
  if (slot < 0) return c->fail(GD_ERR_STATE, "component not monitored");
  const size_t per = c->dft_f.size() * c->dft_plane;
  GD_HIP(c, hipStreamSynchronize(c->stream));
  GD_HIP(c, hipMemcpy(re, c->dft_re + slot * per, per * sizeof(double), hipMemcpyDeviceToHost));
  GD_HIP(c, hipMemcpy(im, c->dft_im + slot * per, per * sizeof(double), hipMemcpyDeviceToHost));
  return GD_OK;
}

// ---- stepping -----------------------------------------------------------------

int gd_plan_tiling(const gd_ctx* c, gd_tiling* out) {
  if (!c || !out) return GD_ERR_ARG;
  *out = plan(c);
  return GD_OK;
}

int gd_plan_tiling_for(size_t nx, size_t ny, size_t nz, int storage, gd_tiling* out) {
  if (!out || nx < 3 || ny < 3 || nz < 3) return GD_ERR_ARG;
  if (storage != GD_STORAGE_F32 && storage != GD_STORAGE_F64 && !is16(storage))
    return GD_ERR_ARG;
  *out = plan_dims(nx, ny, nz, storage);
  return GD_OK;
}

int gd_set_launch_timing(gd_ctx* c, int enable) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  c->launch_timing = enable != 0;
  return GD_OK;
}

int gd_last_run_stats(const gd_ctx* c, double* total_ms, double* kernel_ms, size_t* nlaunch) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(const_cast<gd_ctx*>(c));
  if (total_ms) *total_ms = c->last_total_ms;
  if (kernel_ms) *kernel_ms = c->last_kernel_ms;
  if (nlaunch) *nlaunch = c->last_nlaunch;
  return GD_OK;
}

int gd_device_bytes(const gd_ctx* c, size_t* bytes) {
  if (!c || !bytes) return GD_ERR_ARG;
  size_t b = 0;
  for (int k = 0; k < 2; ++k)
    for (int a = 0; a < 6; ++a)
      if (c->f[k][a]) b += c->N * c->esz;
  for (int a = 0; a < 12; ++a)
    if (c->coeff[a]) b += c->N * (c->storage == GD_STORAGE_F64 ? 8 : 4);
  if (c->cb) b += c->N * (c->storage == GD_STORAGE_F64 ? 8 : 4);
  if (c->mat_id) b += c->N + GD_MAX_MATERIALS * sizeof(float);
  if (c->dft_snap) b += (size_t)c->dft_snap_levels * c->dft_ncomp * c->dft_plane * sizeof(float);
  for (auto& s : c->src) b += s.len * sizeof(float);
  b += 2 * (size_t)c->dft_ncomp * c->dft_f.size() * c->dft_plane * sizeof(double);
  b += c->dft_w_cap * sizeof(double);
  *bytes = b;
  return GD_OK;
}

int gd_run(gd_ctx* c, int variant, size_t nsteps, const gd_tiling* tiling) {
  const int rc = gd_run_async(c, variant, nsteps, tiling);
  if (rc != GD_OK) return rc;
  return settle(c);
}

int gd_sync(gd_ctx* c) {
  if (!c) return GD_ERR_ARG;
  return settle(c);
}

int gd_run_async(gd_ctx* c, int variant, size_t nsteps, const gd_tiling* tiling) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (variant != GD_VARIANT_PLAIN && variant != GD_VARIANT_TILED && variant != GD_VARIANT_FUSED)
    return c->fail(GD_ERR_ARG, "unknown variant");
  if (!c->have_mat) return c->fail(GD_ERR_STATE, "no material set");
  if (c->step + nsteps > (size_t)INT32_MAX) return c->fail(GD_ERR_ARG, "step count too large");
  for (auto& s : c->src)
    if (c->step + nsteps > s.len)
      return c->fail(GD_ERR_ARG, "source table does not cover the requested steps");
  c->last_total_ms = c->last_kernel_ms = 0.0;
  c->last_nlaunch = 0;
  if (nsteps == 0) return GD_OK;

  gd::ZTiling zt{};
  if (variant == GD_VARIANT_TILED) {
    gd_tiling t = plan(c);
    if (tiling) {
      if (tiling->blt) t.blt = tiling->blt;
      if (tiling->rows) t.rows = tiling->rows;
      if (tiling->waves) t.waves = tiling->waves;
      if (tiling->zchunk) t.zchunk = tiling->zchunk;
    }
    zt = gd::ZTiling{t.blt, t.rows, t.waves, t.zchunk};
    // A monitored z plane gets a z chunk of its own (gd::chunk_range): only
    // that one-plane layer of tiles then runs the build with the snapshot
    // code, and the rest of its chunk the lean build. Done for launches of
    // up to 8192 tiles (32 rounds of 256 CUs): measured at 512^3 with 1440
    // (binary16) and 3200 (FP32) tiles +1.9 % and +0.5 %, at 1024^3 binary16
    // with 11520 tiles -0.7 % (many tiles per CU hide the heavier layer among
    // the lean ones, and the cut only adds a z halo).
    if (c->dft_ncomp && c->dft_axis == 2 && gd::zstream_lean_supported(zt)) {
      int ntx = 0, nty = 0, ntz = 0;
      gd::ztile_counts(zt, (int)c->nx, (int)c->ny, (int)c->nz, ntx, nty, ntz);
      if ((long long)ntx * nty * ntz <= 8192) zt.zsplit = (int)c->dft_index;
    }
    if (c->storage == GD_STORAGE_F64)
      return c->fail(GD_ERR_UNSUPPORTED, "the tiled kernel runs FP32 and 16-bit storage only");
    if (c->nx * c->ny * 4 >= (size_t(1) << 31))
      return c->fail(GD_ERR_UNSUPPORTED, "the tiled kernel needs an xy plane below 512 Mi cells");
    if (!gd::ztiling_supported(zt, (gd::Storage)c->storage, c->mat_kind))
      return c->fail(GD_ERR_UNSUPPORTED, "unsupported tiling (blt 1..4, rows x waves in "
                                         "{2,4}x{4,8}, at least one owned row)");
    int rc = alloc_fields(c, 1);
    if (rc != GD_OK) return rc;
    if (c->dft_ncomp && !c->dft_snap) {
      // as many levels as fit in 256 MiB, between kMaxBlt and 32: each fold
      // of the snapshots into the FP64 sums is a kernel and a dependency
      // between two launches, and reads and writes the sums once
      const size_t per_level = (size_t)c->dft_ncomp * c->dft_plane * sizeof(float);
      const size_t fit = (size_t(256) << 20) / std::max<size_t>(1, per_level);
      c->dft_snap_levels = (int)std::min<size_t>(32, std::max<size_t>(gd::kMaxBlt, fit));
      const size_t n = (size_t)c->dft_snap_levels * c->dft_ncomp * c->dft_plane;
      hipError_t e = hipMalloc(&c->dft_snap, n * sizeof(float));
      if (e != hipSuccess) {
        c->dft_snap_levels = 0;
        return c->hip_fail(e, "hipMalloc(dft snapshot)");
      }
    }
  }

  if (variant == GD_VARIANT_FUSED) {
    if (c->storage == GD_STORAGE_F64 && c->mat_kind == gd::kMatIdx)
      return c->fail(GD_ERR_UNSUPPORTED, "FP64 storage takes the dense material");
    if (c->nx * c->ny * 8 >= (size_t(1) << 32))
      return c->fail(GD_ERR_UNSUPPORTED, "the fused kernel needs an xy plane below 512 Mi cells");
    int rc = alloc_fields(c, 1);
    if (rc != GD_OK) return rc;
  }

  gd::Grid g;
  g.nx = (int)c->nx;
  g.ny = (int)c->ny;
  g.nz = (int)c->nz;
  g.sxy = (long long)(c->nx * c->ny);
  g.dx = c->dx;

  gd::Material m{};
  for (int a = 0; a < 12; ++a) m.c[a] = c->coeff[a];
  m.cb = c->cb;
  m.id = c->mat_id;
  m.table = c->mat_table;
  m.db = c->db;
  m.db64 = (double)c->db;
  const gd::Scales sc = scales_of(c);
  const gd::Storage st = (gd::Storage)c->storage;

  gd::Sources s{};
  s.n = (int)c->src.size();
  if (s.n) {
    if (!c->src_desc) {
      hipError_t e = hipMalloc(&c->src_desc, GD_MAX_SOURCES * sizeof(gd::SourceDesc));
      if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(source descriptors)");
    }
    c->src_desc_host.resize(s.n);
    for (int q = 0; q < s.n; ++q) {
      gd::SourceDesc& sd = c->src_desc_host[q];
      sd.comp = c->src[q].comp;
      sd.i = (int)c->src[q].i;
      sd.j = (int)c->src[q].j;
      sd.k = (int)c->src[q].k;
      sd.table = c->src[q].table;
    }
    GD_HIP(c, hipMemcpyAsync(c->src_desc, c->src_desc_host.data(), s.n * sizeof(gd::SourceDesc),
                             hipMemcpyHostToDevice, c->stream));
    s.d = c->src_desc;
  }

  gd::Dft d{};
  d.ncomp = c->dft_ncomp;
  if (d.ncomp) {
    const size_t nf = c->dft_f.size();
    const size_t need = 2 * nf * nsteps;
    // one cos and one sin weight per frequency and step of this call
    if (nf * nsteps > (size_t(1) << 27))
      return c->fail(GD_ERR_ARG, "DFT weight table above 2 GiB: run fewer steps per call");
    if (need > c->dft_w_cap) {
      GD_HIP(c, hipStreamSynchronize(c->stream));
      if (c->dft_w) (void)hipFree(c->dft_w);
      c->dft_w = nullptr;
      c->dft_w_cap = 0;
      hipError_t e = hipMalloc(&c->dft_w, need * sizeof(double));
      if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(dft weights)");
      c->dft_w_cap = need;
    }
    // w = cos(2 pi f n)*scale, sin(2 pi f n)*scale in double (the reference
    // evaluates cos(omega*t) per cell on the device, kernels.cuh:109-110).
    c->dft_w_host.resize(need);
    for (size_t q = 0; q < nf; ++q)
      for (size_t r = 0; r < nsteps; ++r)
        gd::dft_weight(c->dft_f[q], (long long)(c->step + r), c->dft_scale,
                       &c->dft_w_host[q * nsteps + r], &c->dft_w_host[nf * nsteps + q * nsteps + r]);
    GD_HIP(c, hipMemcpyAsync(c->dft_w, c->dft_w_host.data(), need * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    d.mask = 0u;
    for (int q = 0; q < d.ncomp; ++q) d.mask |= 1u << c->dft_comp[q];
    d.axis = c->dft_axis;
    d.index = (int)c->dft_index;
    d.nf = (int)nf;
    d.run_n0 = (int)c->step;
    d.run_len = (int)nsteps;
    d.plane = (long long)c->dft_plane;
    d.wc = c->dft_w;
    d.ws = c->dft_w + nf * nsteps;
    d.re = c->dft_re;
    d.im = c->dft_im;
    d.snap = c->dft_snap;
  }

  size_t nev = 0;
  GD_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (variant == GD_VARIANT_PLAIN) {
    const gd::Fields F = fields_of(c, c->cur);
    for (size_t r = 0; r < nsteps; ++r) {
      hipEvent_t a = nullptr, b = nullptr;
      if (c->launch_timing) {
        a = pool_event(c, nev++);
        b = pool_event(c, nev++);
        if (!a || !b) return c->fail(GD_ERR_HIP, "event creation failed");
        GD_HIP(c, hipEventRecord(a, c->stream));
      }
      GD_HIP(c, gd::launch_plain_step(st, F, sc, m, c->mat_kind, g, s, d, (int)(c->step + r),
                                      (int)r, c->stream));
      if (c->launch_timing) GD_HIP(c, hipEventRecord(b, c->stream));
    }
  } else if (variant == GD_VARIANT_FUSED) {
    for (size_t r = 0; r < nsteps; ++r) {
      hipEvent_t a = nullptr, b = nullptr;
      if (c->launch_timing) {
        a = pool_event(c, nev++);
        b = pool_event(c, nev++);
        if (!a || !b) return c->fail(GD_ERR_HIP, "event creation failed");
        GD_HIP(c, hipEventRecord(a, c->stream));
      }
      GD_HIP(c, gd::launch_fused_step(st, fields_of(c, c->cur), fields_of(c, 1 - c->cur), sc, m,
                                      c->mat_kind, g, s, d, (int)(c->step + r), (int)r,
                                      c->stream));
      if (c->launch_timing) GD_HIP(c, hipEventRecord(b, c->stream));
      c->cur = 1 - c->cur;
    }
  } else {
    // With the planner's 4 x 8 tile a launch is split: the few tiles that hold
    // a source cell in their window or own monitored cells run the full
    // kernel on the side stream while every other tile runs the lean kernel
    // (no source or DFT code in its loop) on the context's stream. In the
    // lean kernel the tiles of a narrow remainder column or a short remainder
    // row share workgroups (zstream_plan_items).
    const bool lean = gd::zstream_lean_supported(zt);
    const bool extras = s.n > 0 || d.ncomp > 0;
    if (lean && extras) {
      if (!c->side) GD_HIP(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
      if (!c->fork_ev) GD_HIP(c, hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
      if (!c->join_ev) GD_HIP(c, hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
    }
    for (auto& sp : c->split) sp.blt = 0;
    // classifies the tiles of tiling t into sp and uploads both id lists on
    // the context's stream (ahead of the launches that read them)
    std::vector<int> lean_ids;
    auto classify = [&](const gd::ZTiling& t, gd_ctx::Split& sp) -> int {
      gd::zstream_plan_items(t, st, g, c->src_desc_host.data(), s.n, d, sp.ids,
                             lean_ids, sp.packs);
      sp.nextra = sp.ids.size();
      sp.ids.insert(sp.ids.end(), lean_ids.begin(), lean_ids.end());
      const size_t n = sp.ids.size();
      if (n > sp.cap) {
        if (sp.d_ids) (void)hipFree(sp.d_ids);
        sp.d_ids = nullptr;
        sp.cap = 0;
        hipError_t e = hipMalloc(&sp.d_ids, n * sizeof(int));
        if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(tile lists)");
        sp.cap = n;
      }
      GD_HIP(c, hipMemcpyAsync(sp.d_ids, sp.ids.data(), n * sizeof(int), hipMemcpyHostToDevice,
                               c->stream));
      const size_t np = sp.packs.size();
      if (np > sp.pcap) {
        if (sp.d_packs) (void)hipFree(sp.d_packs);
        sp.d_packs = nullptr;
        sp.pcap = 0;
        hipError_t e = hipMalloc(&sp.d_packs, np * sizeof(gd::ZPack));
        if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(tile packs)");
        sp.pcap = np;
      }
      if (np)
        GD_HIP(c, hipMemcpyAsync(sp.d_packs, sp.packs.data(), np * sizeof(gd::ZPack),
                                 hipMemcpyHostToDevice, c->stream));
      sp.blt = t.blt;
      return GD_OK;
    };
    // the snapshot buffer holds dft_snap_levels levels (32 for the 512^3 and
    // 1024^3 workloads): the DFT fold runs once per 8 to 10 launches, not
    // after every one
    size_t r = 0;
    int snap_levels = 0;  // levels in the snapshot buffer not folded yet
    size_t snap_r0 = 0;   // run step of the first of them
    while (r < nsteps) {
      gd::ZTiling t = zt;
      t.blt = (int)std::min<size_t>((size_t)zt.blt, nsteps - r);
      auto launch = [&](const gd::ZPart& part, hipStream_t strm) {
        return gd::launch_zstream(t, st, fields_of(c, c->cur), fields_of(c, 1 - c->cur), sc, m,
                                  c->mat_kind, g, s, d, (int)(c->step + r), (int)r, part, strm);
      };
      gd_ctx::Split* sp = nullptr;
      if (lean) {
        sp = &c->split[t.blt == zt.blt ? 0 : 1];
        if (sp->blt != t.blt) {
          const int rc = classify(t, *sp);
          if (rc != GD_OK) return rc;
        }
      }
      hipEvent_t a = nullptr, b = nullptr;
      if (c->launch_timing) {
        a = pool_event(c, nev++);
        b = pool_event(c, nev++);
        if (!a || !b) return c->fail(GD_ERR_HIP, "event creation failed");
        GD_HIP(c, hipEventRecord(a, c->stream));
      }
      d.snap_base = snap_levels;
      const int ne = sp ? (int)sp->nextra : 0;
      const int nl = sp ? (int)(sp->ids.size() - sp->nextra) : 0;
      if (!sp || nl == 0) {
        GD_HIP(c, launch(gd::ZPart{gd::kZFull, nullptr, 0}, c->stream));
      } else if (ne == 0) {
        GD_HIP(c, launch(gd::ZPart{gd::kZLean, sp->d_ids, nl, sp->d_packs}, c->stream));
      } else {
        GD_HIP(c, hipEventRecord(c->fork_ev, c->stream));
        GD_HIP(c, hipStreamWaitEvent(c->side, c->fork_ev, 0));
        GD_HIP(c, launch(gd::ZPart{gd::kZFull, sp->d_ids, ne}, c->side));
        GD_HIP(c, launch(gd::ZPart{gd::kZLean, sp->d_ids + ne, nl, sp->d_packs}, c->stream));
        GD_HIP(c, hipEventRecord(c->join_ev, c->side));
        GD_HIP(c, hipStreamWaitEvent(c->stream, c->join_ev, 0));
      }
      if (c->launch_timing) GD_HIP(c, hipEventRecord(b, c->stream));
      if (d.ncomp) {
        snap_levels += t.blt;
        const size_t rn = r + (size_t)t.blt;
        const size_t next = std::min<size_t>((size_t)zt.blt, nsteps - rn);
        if (rn >= nsteps || snap_levels + (int)next > c->dft_snap_levels) {
          GD_HIP(c, gd::launch_dft_snapshot(d, snap_levels, (int)snap_r0, c->stream));
          snap_levels = 0;
          snap_r0 = rn;
        }
      }
      c->cur = 1 - c->cur;
      r += (size_t)t.blt;
    }
  }
  GD_HIP(c, hipEventRecord(c->ev1, c->stream));
  c->pending = true;
  c->pending_nev = nev;
  c->pending_nsteps = nsteps;
  c->pending_variant = variant;
  c->step += nsteps;
  return GD_OK;
}

int gd_calibration_copy(gd_ctx* c, size_t n, int reps, double* ms) {
  if (!c || n == 0 || reps < 1) return GD_ERR_ARG;
  if (c->cal_n != n) {
    for (int k = 0; k < 2; ++k) {
      if (c->cal[k]) (void)hipFree(c->cal[k]);
      c->cal[k] = nullptr;
      hipError_t e = hipMalloc(&c->cal[k], n * sizeof(float));
      if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(calibration)");
      GD_HIP(c, hipMemsetAsync(c->cal[k], 0, n * sizeof(float), c->stream));
    }
    c->cal_n = n;
  }
  GD_HIP(c, hipEventRecord(c->ev0, c->stream));
  for (int q = 0; q < reps; ++q) GD_HIP(c, gd::launch_copy(c->cal[0], c->cal[1], n, c->stream));
  GD_HIP(c, hipEventRecord(c->ev1, c->stream));
  GD_HIP(c, hipStreamSynchronize(c->stream));
  float t = 0.0f;
  GD_HIP(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
  if (ms) *ms = t / reps;
  return GD_OK;
}

// ---- learned coarse-grid corrector -------------------------------------------

int gd_restrict_fields(gd_ctx* c, float* const out[6]) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  GD_SETTLE(c);
  if (!out) return c->fail(GD_ERR_ARG, "null output table");
  for (int a = 0; a < 6; ++a)
    if (!out[a]) return c->fail(GD_ERR_ARG, "null output array");
  if (c->storage != GD_STORAGE_F32) return c->fail(GD_ERR_UNSUPPORTED, "FP32 storage only");
  const size_t cn = (c->nx / 2) * (c->ny / 2) * (c->nz / 2);
  if (cn == 0) return c->fail(GD_ERR_ARG, "grid too small to restrict");
  if (!c->restrict_buf) {
    hipError_t e = hipMalloc(&c->restrict_buf, 6 * cn * sizeof(float));
    if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(restriction)");
  }
  GD_HIP(c, gd::launch_restrict(fields_of(c, c->cur), (int)c->nx, (int)c->ny, (int)c->nz,
                                c->restrict_buf, c->stream));
  GD_HIP(c, hipStreamSynchronize(c->stream));
  for (int a = 0; a < 6; ++a)
    GD_HIP(c, hipMemcpy(out[a], c->restrict_buf + a * cn, cn * sizeof(float),
                        hipMemcpyDeviceToHost));
  return GD_OK;
}

int gd_run_corrected(gd_ctx* c, gd_nn* nn, size_t nsteps, int interval, const float cfg[6],
                     const gd_tiling* tiling) {
  if (!c) return GD_ERR_ARG;
  GD_SETTLE(c);
  if (!nn || !cfg) return c->fail(GD_ERR_ARG, "null network or configuration");
  if (interval < 2) return c->fail(GD_ERR_ARG, "interval must be >= 2");
  if (c->storage != GD_STORAGE_F32 || !c->have_mat || c->mat_kind == gd::kMatFull)
    return c->fail(GD_ERR_UNSUPPORTED, "the corrector needs FP32 fields and the isotropic "
                                       "material");
  int nx, ny, nz;
  gd::nn_dims(nn, &nx, &ny, &nz);
  if ((size_t)nx != c->nx || (size_t)ny != c->ny || (size_t)nz != c->nz)
    return c->fail(GD_ERR_ARG, "network grid differs from the context grid");
  for (int q = 0; q < 5; ++q)
    if (!(cfg[q] > 0.0f)) return c->fail(GD_ERR_ARG, "scales must be positive");
  for (auto& p : c->prev_e)
    if (!p) {
      hipError_t e = hipMalloc(&p, c->N * sizeof(float));
      if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(previous E)");
    }
  const float* cb = static_cast<const float*>(c->cb);
  if (c->mat_kind == gd::kMatIdx) {
    if (!c->cb_dense) {
      hipError_t e = hipMalloc(&c->cb_dense, c->N * sizeof(float));
      if (e != hipSuccess) return c->hip_fail(e, "hipMalloc(dense Cb)");
    }
    GD_HIP(c, gd::launch_expand_material(c->mat_id, c->mat_table, c->N, c->cb_dense, c->stream));
    cb = c->cb_dense;
  }
  gd::NnFeat nf{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5]};
  size_t done = 0;
  while (done + (size_t)interval <= nsteps) {
    int rc = gd_run(c, GD_VARIANT_TILED, (size_t)interval - 1, tiling);
    if (rc != GD_OK) return rc;
    for (int a = 0; a < 3; ++a)
      GD_HIP(c, hipMemcpyAsync(c->prev_e[a], c->f[c->cur][a], c->N * sizeof(float),
                               hipMemcpyDeviceToDevice, c->stream));
    rc = gd_run(c, GD_VARIANT_TILED, 1, tiling);
    if (rc != GD_OK) return rc;
    GD_HIP(c, gd::nn_correct(nn, fields_of(c, c->cur), c->prev_e[0], c->prev_e[1], c->prev_e[2],
                             cb, nf, c->stream));
    GD_HIP(c, hipStreamSynchronize(c->stream));
    done += (size_t)interval;
  }
  if (done < nsteps) return gd_run(c, GD_VARIANT_TILED, nsteps - done, tiling);
  return GD_OK;
}

}  // extern "C"

namespace {

int settle(gd_ctx* c) {
  if (!c->pending) return GD_OK;
  c->pending = false;
  GD_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0.0f;
  GD_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_total_ms = ms;
  if (c->launch_timing) {
    const size_t nev = c->pending_nev;
    double sum = 0.0;
    for (size_t q = 0; q + 1 < nev; q += 2) {
      float t = 0.0f;
      GD_HIP(c, hipEventElapsedTime(&t, c->ev_pool[q], c->ev_pool[q + 1]));
      sum += t;
    }
    c->last_kernel_ms = sum;
    // PLAIN: an E and an H launch per step
    c->last_nlaunch = c->pending_variant == GD_VARIANT_PLAIN ? c->pending_nsteps * 2 : nev / 2;
  }
  return GD_OK;
}

}  // namespace
