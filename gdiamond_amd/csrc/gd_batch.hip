// Batches of independent FDTD problems: one persistent TILED launch per blt
// steps covers every instance of a launch group, launch groups run as
// parallel branches, and a run can be captured once in a hipGraph and
// replayed (the absolute step and each group's field-copy parity live in
// device memory, so replays need no new arguments).
//
// No counterpart in the reference: it runs one problem per process and fixes
// tile shapes at compile time (kernels_mm_ver2.cuh:8-32, README.md:32-33).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/gdiamond/gd_capi.h"
#include "gd_common.hpp"

namespace {

struct Inst {
  size_t nx = 0, ny = 0, nz = 0, N = 0;
  int cls = 0;
  float* f[2][6] = {{nullptr}};  // field copies (in the arena)
  float* cb = nullptr;           // dense Cb (in the arena)
  uint8_t* id = nullptr;         // material id per cell (in the arena)
  float* table = nullptr;        // Cb per id, kMaxMaterials entries (in the arena)
  bool indexed = false;          // id/table hold the material (<= kMaxMaterials values)
  float db = 0.0f;
  bool have_mat = false;
  struct Src {
    int comp;
    size_t i, j, k;
    float* table;
    size_t len;
  };
  std::vector<Src> src;
  int parity = 0;  // field copy holding the current state
};

struct SizeClass {
  size_t nx, ny, nz;
  std::vector<int> insts;
  gd_tiling t;
  bool tuned = false;
};

// One launch configuration (a tiling) over a set of instances: per-instance
// launch arguments for both copy parities and the work items, instance by
// instance (largest first), each instance's tiles in tile order. Items whose
// tile holds a source cell are flagged: the batch kernel runs them with the
// point-source code, all others with the build without it.
struct Plan {
  gd::ZTiling t{};
  gd::ZArgs* d_args = nullptr;   // [2 * ninst]
  gd::BatchItem* d_items = nullptr;
  int nitems = 0;
  int nfull = 0;    // flagged items
  int nblocks = 0;  // workgroups of a launch
  std::vector<int> insts;
};

// Queue counters of one launch.
constexpr int kCountersPerLaunch = 1;

struct Group {
  int blt = 0, rows = 0, waves = 0;
  std::vector<int> insts;      // largest first
  Plan main;
  std::map<int, Plan> rem;     // remainder launches of fewer levels, by blt
};

// Per-nsteps schedule: queue counters of every launch, the parity flips of
// the groups and (in graph mode) the captured graph.
struct Schedule {
  unsigned* d_counters = nullptr;
  int ncounters = 0;
  int* d_flip = nullptr;
  std::vector<int> nlaunch;  // per group
  hipGraphExec_t exec = nullptr;
};

gd_tiling default_tiling(size_t nx, size_t ny, size_t nz) {
  gd_tiling t;
  t.blt = 4;
  t.rows = 4;
  t.waves = 8;
  const size_t ox = 64 - 2 * t.blt, oy = (size_t)(t.rows * t.waves - 2 * t.blt);
  const size_t txy = ((nx + ox - 1) / ox) * ((ny + oy - 1) / oy);
  size_t zc = nz * txy / 1024;
  zc = std::max<size_t>(16, std::min<size_t>(64, zc));
  t.zchunk = (int)std::min(zc, nz);
  return t;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Arrays of one instance start 16 KiB + 256 B past the 256-byte boundary
// after the previous one, so that arrays a power of two in size (32^3, 64^3,
// 128^3 FP32) do not put the same cell of every component on the same
// memory channel and bank pattern.
constexpr size_t kArrayStagger = 16640;
size_t array_stride(size_t N) { return align_up(N * sizeof(float), 256) + kArrayStagger; }

// Arena bytes of one instance: 12 field arrays and the dense Cb (FP32), the
// material ids (1 byte per cell) and the id table.
size_t inst_bytes(size_t N) {
  return 13 * array_stride(N) + align_up(N, 256) +
         align_up(gd::kMaxMaterials * sizeof(float), 256);
}

}  // namespace

struct gd_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> gstreams;
  std::vector<hipEvent_t> join_ev;
  hipEvent_t fork_ev = nullptr, ev0 = nullptr, ev1 = nullptr;
  float dx = 1.0f;
  std::vector<Inst> inst;
  std::vector<SizeClass> cls;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  gd::SourceDesc* d_srcdesc = nullptr;  // [ninst * kMaxSources]
  bool src_dirty = true;

  size_t step = 0;
  int* d_step = nullptr;
  int* d_parity = nullptr;  // per group (capacity kMaxGroups)
  static constexpr int kMaxGroups = 64;

  bool plan_valid = false;
  std::vector<Group> groups;
  std::map<size_t, Schedule> sched;

  double last_ms = 0.0;
  size_t last_nlaunch = 0;
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

#define BH(b, call)                                        \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return (b)->hip_fail(e_, #call); \
  } while (0)

namespace {

thread_local std::string g_batch_create_error;

void free_plan(Plan& p) {
  if (p.d_args) (void)hipFree(p.d_args);
  if (p.d_items) (void)hipFree(p.d_items);
  p = Plan();
}

void free_schedules(gd_batch* b) {
  for (auto& kv : b->sched) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.d_counters) (void)hipFree(kv.second.d_counters);
    if (kv.second.d_flip) (void)hipFree(kv.second.d_flip);
  }
  b->sched.clear();
}

void invalidate(gd_batch* b) {
  if (b->plan_valid) (void)hipStreamSynchronize(b->stream);
  for (auto& g : b->groups) {
    free_plan(g.main);
    for (auto& kv : g.rem) free_plan(kv.second);
  }
  b->groups.clear();
  free_schedules(b);
  b->plan_valid = false;
}

// Material form of the batch's launches: the indexed one when every instance
// has it, else the dense one.
gd::MatKind batch_mat(const gd_batch* b) {
  for (const auto& in : b->inst)
    if (!in.indexed) return gd::kMatIso;
  return gd::kMatIdx;
}

gd::ZArgs make_args(const gd_batch* b, int i, int parity, const gd::ZTiling& t) {
  const Inst& in = b->inst[i];
  gd::ZArgs a{};
  for (int c = 0; c < 6; ++c) {
    a.src.c[c] = in.f[parity][c];
    a.dst.c[c] = in.f[1 - parity][c];
  }
  for (int c = 0; c < 6; ++c) {
    a.sc.s[c] = 1.0f;
    a.sc.inv[c] = 1.0f;
  }
  a.m.cb = in.cb;
  a.m.id = in.id;
  a.m.table = in.table;
  a.m.db = in.db;
  a.m.db64 = (double)in.db;
  a.g.nx = (int)in.nx;
  a.g.ny = (int)in.ny;
  a.g.nz = (int)in.nz;
  a.g.sxy = (long long)(in.nx * in.ny);
  a.g.dx = b->dx;
  a.s.n = (int)in.src.size();
  a.s.d = b->d_srcdesc + (size_t)i * gd::kMaxSources;
  a.d.ncomp = 0;
  a.n0 = 0;
  a.run_step0 = 0;
  gd::ztile_counts(t, a.g.nx, a.g.ny, a.g.nz, a.ntx, a.nty, a.ntz);
  a.zc = t.zchunk;
  a.zsplit = -1;
  return a;
}

// Builds a plan for `insts` (already ordered, largest first); the tiling's
// zchunk is replaced per instance by its class's zchunk when per_class.
int build_plan(gd_batch* b, const std::vector<int>& insts, const gd::ZTiling& t, bool per_class,
               Plan& p) {
  p.t = t;
  p.insts = insts;
  const size_t ninst = b->inst.size();
  std::vector<gd::ZArgs> args(2 * ninst);
  std::memset(args.data(), 0, args.size() * sizeof(gd::ZArgs));
  std::vector<gd::BatchItem> items;
  std::vector<int> extra_ids, lean_ids;
  std::vector<gd::SourceDesc> sd;
  gd::Dft no_dft{};
  no_dft.ncomp = 0;
  for (int i : insts) {
    gd::ZTiling ti = t;
    if (per_class) ti.zchunk = b->cls[b->inst[i].cls].t.zchunk;
    for (int par = 0; par < 2; ++par) args[2 * i + par] = make_args(b, i, par, ti);
    const gd::ZArgs& a = args[2 * i];
    const long long nt = (long long)a.ntx * a.nty * a.ntz;
    if (nt > (1ll << 30)) return b->fail(GD_ERR_ARG, "too many tiles");
    // tiles with a source cell in their window run with the source code (the
    // batch tile order is x fastest, then y, then z: ids are positions, and
    // both lists come in that order)
    sd.clear();
    for (const auto& s : b->inst[i].src)
      sd.push_back(gd::SourceDesc{s.comp, (int)s.i, (int)s.j, (int)s.k, s.table});
    gd::zstream_classify(ti, 0, a.g, sd.data(), (int)sd.size(), no_dft, extra_ids, lean_ids);
    size_t qe = 0, ql = 0;
    while (qe < extra_ids.size() || ql < lean_ids.size()) {
      const bool e = ql == lean_ids.size() ||
                     (qe < extra_ids.size() && extra_ids[qe] < lean_ids[ql]);
      items.push_back(e ? gd::BatchItem{i, extra_ids[qe++], 1}
                        : gd::BatchItem{i, lean_ids[ql++], 0});
    }
    p.nfull += (int)extra_ids.size();
  }
  if (items.size() > (size_t)INT32_MAX) return b->fail(GD_ERR_ARG, "too many work items");
  BH(b, hipMalloc(&p.d_args, args.size() * sizeof(gd::ZArgs)));
  BH(b, hipMalloc(&p.d_items, std::max<size_t>(1, items.size()) * sizeof(gd::BatchItem)));
  BH(b, hipMemcpy(p.d_args, args.data(), args.size() * sizeof(gd::ZArgs), hipMemcpyHostToDevice));
  if (!items.empty())
    BH(b, hipMemcpy(p.d_items, items.data(), items.size() * sizeof(gd::BatchItem),
                    hipMemcpyHostToDevice));
  p.nitems = (int)items.size();
  p.nblocks = gd::zstream_batch_blocks(t, batch_mat(b), b->device);
  if (p.nblocks < 1) return b->fail(GD_ERR_HIP, "occupancy query failed");
  return GD_OK;
}

int upload_sources(gd_batch* b) {
  if (!b->src_dirty) return GD_OK;
  std::vector<gd::SourceDesc> h(b->inst.size() * gd::kMaxSources);
  std::memset(h.data(), 0, h.size() * sizeof(gd::SourceDesc));
  for (size_t i = 0; i < b->inst.size(); ++i)
    for (size_t q = 0; q < b->inst[i].src.size(); ++q) {
      const auto& s = b->inst[i].src[q];
      h[i * gd::kMaxSources + q] = gd::SourceDesc{s.comp, (int)s.i, (int)s.j, (int)s.k, s.table};
    }
  BH(b, hipMemcpy(b->d_srcdesc, h.data(), h.size() * sizeof(gd::SourceDesc),
                  hipMemcpyHostToDevice));
  b->src_dirty = false;
  return GD_OK;
}

// Every instance's current state into copy 0 (so that the instances of any
// group share one parity), device parities to 0.
int normalize_parity(gd_batch* b) {
  for (auto& in : b->inst) {
    if (in.parity == 0) continue;
    for (int c = 0; c < 6; ++c)
      BH(b, hipMemcpyAsync(in.f[0][c], in.f[1][c], in.N * sizeof(float), hipMemcpyDeviceToDevice,
                           b->stream));
    in.parity = 0;
  }
  BH(b, hipMemsetAsync(b->d_parity, 0, gd_batch::kMaxGroups * sizeof(int), b->stream));
  BH(b, hipStreamSynchronize(b->stream));
  return GD_OK;
}

int ensure_plan(gd_batch* b) {
  int rc = upload_sources(b);
  if (rc != GD_OK) return rc;
  if (b->plan_valid) return GD_OK;
  rc = normalize_parity(b);
  if (rc != GD_OK) return rc;
  // groups: classes with equal (blt, rows, waves)
  std::vector<Group> gs;
  for (size_t c = 0; c < b->cls.size(); ++c) {
    const gd_tiling& t = b->cls[c].t;
    Group* g = nullptr;
    for (auto& x : gs)
      if (x.blt == t.blt && x.rows == t.rows && x.waves == t.waves) g = &x;
    if (!g) {
      gs.push_back(Group());
      g = &gs.back();
      g->blt = t.blt;
      g->rows = t.rows;
      g->waves = t.waves;
    }
    for (int i : b->cls[c].insts) g->insts.push_back(i);
  }
  if ((int)gs.size() > gd_batch::kMaxGroups) return b->fail(GD_ERR_UNSUPPORTED, "too many groups");
  for (auto& g : gs) {
    std::stable_sort(g.insts.begin(), g.insts.end(),
                     [&](int x, int y) { return b->inst[x].N > b->inst[y].N; });
    // the zchunk of a group plan is per instance (its class's); 1 is a placeholder
    gd::ZTiling t{g.blt, g.rows, g.waves, 1};
    rc = build_plan(b, g.insts, t, true, g.main);
    if (rc != GD_OK) {
      b->groups = std::move(gs);
      invalidate(b);
      return rc;
    }
  }
  b->groups = std::move(gs);
  b->plan_valid = true;
  return GD_OK;
}

Plan* plan_for(gd_batch* b, Group& g, int blt, int* rc) {
  *rc = GD_OK;
  if (blt == g.blt) return &g.main;
  auto it = g.rem.find(blt);
  if (it != g.rem.end()) return &it->second;
  Plan& p = g.rem[blt];
  gd::ZTiling t{blt, g.rows, g.waves, 1};
  *rc = build_plan(b, g.insts, t, true, p);
  return *rc == GD_OK ? &p : nullptr;
}

int get_schedule(gd_batch* b, size_t nsteps, Schedule** out) {
  auto it = b->sched.find(nsteps);
  if (it != b->sched.end()) {
    *out = &it->second;
    return GD_OK;
  }
  Schedule s;
  std::vector<int> flip;
  for (auto& g : b->groups) {
    const int L = (int)((nsteps + g.blt - 1) / g.blt);
    s.nlaunch.push_back(L);
    s.ncounters += kCountersPerLaunch * L;
    flip.push_back(L & 1);
    if (nsteps % g.blt) {
      int rc;
      if (!plan_for(b, g, (int)(nsteps % g.blt), &rc)) return rc;
    }
  }
  BH(b, hipMalloc(&s.d_counters, std::max(1, s.ncounters) * sizeof(unsigned)));
  BH(b, hipMemset(s.d_counters, 0, std::max(1, s.ncounters) * sizeof(unsigned)));
  BH(b, hipMalloc(&s.d_flip, std::max<size_t>(1, flip.size()) * sizeof(int)));
  if (!flip.empty())
    BH(b, hipMemcpy(s.d_flip, flip.data(), flip.size() * sizeof(int), hipMemcpyHostToDevice));
  b->sched[nsteps] = s;
  *out = &b->sched[nsteps];
  return GD_OK;
}

int ensure_streams(gd_batch* b, size_t n) {
  while (b->gstreams.size() < n) {
    hipStream_t s;
    hipEvent_t e;
    BH(b, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    BH(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->gstreams.push_back(s);
    b->join_ev.push_back(e);
  }
  return GD_OK;
}

// One launch of plan p on `stream`: one persistent kernel over every item,
// with its queue counter. L carries the other launch fields; nk counts the
// kernels issued.
int launch_plan(gd_batch* b, const Plan& p, gd::BatchLaunch L, unsigned* counter,
                hipStream_t stream, size_t* nk) {
  if (p.nitems < 1) return GD_OK;
  L.inst = p.d_args;
  L.items = p.d_items;
  L.nitems = p.nitems;
  L.counter = counter;
  BH(b, gd::launch_zstream_batch(p.t, batch_mat(b), L, p.nblocks, stream));
  ++*nk;
  return GD_OK;
}

// Issues the persistent launches of one run (every group on its own stream,
// forked from and joined back into the main stream) and the advance kernel.
int issue_run(gd_batch* b, size_t nsteps, Schedule& s, size_t* nlaunch) {
  BH(b, hipEventRecord(b->fork_ev, b->stream));
  int coff = 0;
  size_t nl = 0;
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    Group& g = b->groups[gi];
    hipStream_t gs = b->gstreams[gi];
    BH(b, hipStreamWaitEvent(gs, b->fork_ev, 0));
    const int L = s.nlaunch[gi];
    for (int k = 0; k < L; ++k) {
      const int blt = (int)std::min<size_t>((size_t)g.blt, nsteps - (size_t)k * g.blt);
      int rc;
      Plan* p = plan_for(b, g, blt, &rc);
      if (!p) return rc;
      gd::BatchLaunch Lh{};
      Lh.step = b->d_step;
      Lh.parity = b->d_parity + gi;
      Lh.step_offset = k * g.blt;
      Lh.launch_index = k;
      size_t nk = 0;
      rc = launch_plan(b, *p, Lh, s.d_counters + coff + kCountersPerLaunch * k, gs, &nk);
      if (rc != GD_OK) return rc;
      ++nl;
    }
    coff += kCountersPerLaunch * L;
    BH(b, hipEventRecord(b->join_ev[gi], gs));
    BH(b, hipStreamWaitEvent(b->stream, b->join_ev[gi], 0));
  }
  BH(b, gd::launch_batch_advance(b->d_step, (int)nsteps, b->d_parity, s.d_flip,
                                 (int)b->groups.size(), s.d_counters, s.ncounters, b->stream));
  *nlaunch = nl + 1;
  return GD_OK;
}

int check_instance(gd_batch* b, int i) {
  if (i < 0 || i >= (int)b->inst.size()) return b->fail(GD_ERR_ARG, "instance out of range");
  return GD_OK;
}

}  // namespace

extern "C" {

int gd_batch_create(gd_batch** out, int ninst, const size_t* dims, float dx) {
  if (!out) return GD_ERR_ARG;
  *out = nullptr;
  if (ninst < 1 || !dims) {
    g_batch_create_error = "need at least one instance";
    return GD_ERR_ARG;
  }
  for (int i = 0; i < ninst; ++i) {
    const size_t nx = dims[3 * i], ny = dims[3 * i + 1], nz = dims[3 * i + 2];
    if (nx < 3 || ny < 3 || nz < 3 || nx > (1u << 20) || ny > (1u << 20) || nz > (1u << 20) ||
        nx * ny * 4 >= (1ull << 31) || nz > (1ull << 36) / (nx * ny)) {
      g_batch_create_error = "instance dims must be >= 3, a plane below 2 GiB and at most 2^36 cells";
      return GD_ERR_ARG;
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_batch_create_error = "no HIP device visible";
    return GD_ERR_NODEVICE;
  }
  gd_batch* b = new gd_batch();
  b->dx = dx;
  b->inst.resize(ninst);
  size_t bytes = 0;
  for (int i = 0; i < ninst; ++i) {
    Inst& in = b->inst[i];
    in.nx = dims[3 * i];
    in.ny = dims[3 * i + 1];
    in.nz = dims[3 * i + 2];
    in.N = in.nx * in.ny * in.nz;
    bytes += inst_bytes(in.N);
    if (bytes > (size_t(1) << 50)) {  // far above any device; also keeps the sum from wrapping
      g_batch_create_error = "batch larger than 1 PiB";
      delete b;
      return GD_ERR_NOMEM;
    }
    int c = -1;
    for (size_t q = 0; q < b->cls.size(); ++q)
      if (b->cls[q].nx == in.nx && b->cls[q].ny == in.ny && b->cls[q].nz == in.nz) c = (int)q;
    if (c < 0) {
      SizeClass sc;
      sc.nx = in.nx;
      sc.ny = in.ny;
      sc.nz = in.nz;
      sc.t = default_tiling(in.nx, in.ny, in.nz);
      b->cls.push_back(sc);
      c = (int)b->cls.size() - 1;
    }
    in.cls = c;
    b->cls[c].insts.push_back(i);
  }
  auto bail = [&](const char* m) {
    g_batch_create_error = m;
    gd_batch_destroy(b);
    return GD_ERR_NOMEM;
  };
  if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess)
    return bail("HIP stream/event creation failed");
  if (hipMalloc(&b->arena, bytes) != hipSuccess) return bail("hipMalloc(batch arena) failed");
  b->arena_bytes = bytes;
  if (hipMemset(b->arena, 0, bytes) != hipSuccess) return bail("hipMemset(batch arena) failed");
  char* p = static_cast<char*>(b->arena);
  for (auto& in : b->inst) {
    const size_t a = array_stride(in.N);
    for (int k = 0; k < 2; ++k)
      for (int c = 0; c < 6; ++c, p += a) in.f[k][c] = reinterpret_cast<float*>(p);
    in.cb = reinterpret_cast<float*>(p);
    p += a;
    in.id = reinterpret_cast<uint8_t*>(p);
    p += align_up(in.N, 256);
    in.table = reinterpret_cast<float*>(p);
    p += align_up(gd::kMaxMaterials * sizeof(float), 256);
  }
  if (hipMalloc(&b->d_srcdesc, (size_t)ninst * gd::kMaxSources * sizeof(gd::SourceDesc)) !=
          hipSuccess ||
      hipMalloc(&b->d_step, sizeof(int)) != hipSuccess ||
      hipMalloc(&b->d_parity, gd_batch::kMaxGroups * sizeof(int)) != hipSuccess)
    return bail("hipMalloc(batch state) failed");
  if (hipMemset(b->d_step, 0, sizeof(int)) != hipSuccess ||
      hipMemset(b->d_parity, 0, gd_batch::kMaxGroups * sizeof(int)) != hipSuccess)
    return bail("hipMemset(batch state) failed");
  *out = b;
  return GD_OK;
}

int gd_batch_destroy(gd_batch* b) {
  if (!b) return GD_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto s : b->gstreams) (void)hipStreamSynchronize(s);
  invalidate(b);
  for (auto& in : b->inst)
    for (auto& s : in.src) (void)hipFree(s.table);
  if (b->arena) (void)hipFree(b->arena);
  if (b->d_srcdesc) (void)hipFree(b->d_srcdesc);
  if (b->d_step) (void)hipFree(b->d_step);
  if (b->d_parity) (void)hipFree(b->d_parity);
  for (auto s : b->gstreams) (void)hipStreamDestroy(s);
  for (auto e : b->join_ev) (void)hipEventDestroy(e);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GD_OK;
}

const char* gd_batch_last_error(const gd_batch* b) {
  return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int gd_batch_set_isotropic(gd_batch* b, int i, const float* cb, float db) {
  if (!b) return GD_ERR_ARG;
  int rc = check_instance(b, i);
  if (rc != GD_OK) return rc;
  if (!cb) return b->fail(GD_ERR_ARG, "null cb array");
  BH(b, hipStreamSynchronize(b->stream));
  Inst& in = b->inst[i];
  BH(b, hipMemcpy(in.cb, cb, in.N * sizeof(float), hipMemcpyHostToDevice));
  // the indexed form as well when the medium has few enough values: a
  // launch runs it when every instance has it (1 byte of material per cell
  // instead of 4; bit-identical results)
  gd::HostVec<uint8_t> id;
  std::vector<float> table;
  const bool indexed = gd::index_materials(cb, in.N, id, table);
  if (indexed) {
    table.resize(gd::kMaxMaterials, 0.0f);
    BH(b, hipMemcpy(in.id, id.data(), in.N, hipMemcpyHostToDevice));
    BH(b, hipMemcpy(in.table, table.data(), gd::kMaxMaterials * sizeof(float),
                    hipMemcpyHostToDevice));
  }
  // db lives in the launch arguments, the material form in the kernel choice
  if (in.db != db || in.indexed != indexed) invalidate(b);
  in.indexed = indexed;
  in.db = db;
  in.have_mat = true;
  return GD_OK;
}

int gd_batch_set_isotropic_eps(gd_batch* b, int i, const float* eps_r, float dt) {
  if (!b) return GD_ERR_ARG;
  int rc = check_instance(b, i);
  if (rc != GD_OK) return rc;
  if (!eps_r) return b->fail(GD_ERR_ARG, "null eps array");
  const Inst& in = b->inst[i];
  std::vector<float> cb(in.N);
  const double ddt = dt, ddx = b->dx;
  for (size_t q = 0; q < in.N; ++q) {
    if (!(eps_r[q] > 0.0f)) return b->fail(GD_ERR_ARG, "eps_r must be positive");
    cb[q] = (float)(ddt / ((double)eps_r[q] * ddx));
  }
  return gd_batch_set_isotropic(b, i, cb.data(), (float)(ddt / ddx));
}

int gd_batch_add_source(gd_batch* b, int i, int comp, size_t si, size_t sj, size_t sk,
                        const float* table, size_t nsteps) {
  if (!b) return GD_ERR_ARG;
  int rc = check_instance(b, i);
  if (rc != GD_OK) return rc;
  Inst& in = b->inst[i];
  if (comp < 0 || comp > 5) return b->fail(GD_ERR_ARG, "component must be 0..5");
  if (si >= in.nx || sj >= in.ny || sk >= in.nz) return b->fail(GD_ERR_ARG, "source outside grid");
  if (!table || nsteps == 0 || nsteps > (size_t)INT32_MAX)
    return b->fail(GD_ERR_ARG, "empty or too long source table");
  if ((int)in.src.size() >= gd::kMaxSources) return b->fail(GD_ERR_ARG, "too many sources");
  for (const auto& o : in.src)
    if (o.comp == comp && o.i == si && o.j == sj && o.k == sk)
      return b->fail(GD_ERR_ARG, "a source on this cell and component already exists");
  Inst::Src s{comp, si, sj, sk, nullptr, nsteps};
  BH(b, hipMalloc(&s.table, nsteps * sizeof(float)));
  hipError_t e = hipMemcpy(s.table, table, nsteps * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(s.table);
    return b->hip_fail(e, "hipMemcpy(source)");
  }
  invalidate(b);  // source counts live in the launch arguments
  in.src.push_back(s);
  b->src_dirty = true;
  return GD_OK;
}

int gd_batch_set_fields(gd_batch* b, int i, const float* const fields[6]) {
  if (!b) return GD_ERR_ARG;
  int rc = check_instance(b, i);
  if (rc != GD_OK) return rc;
  if (!fields) return b->fail(GD_ERR_ARG, "null field table");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return b->fail(GD_ERR_ARG, "null field array");
  BH(b, hipStreamSynchronize(b->stream));
  Inst& in = b->inst[i];
  for (int c = 0; c < 6; ++c)
    BH(b, hipMemcpy(in.f[in.parity][c], fields[c], in.N * sizeof(float), hipMemcpyHostToDevice));
  return GD_OK;
}

int gd_batch_get_fields(gd_batch* b, int i, float* const fields[6]) {
  if (!b) return GD_ERR_ARG;
  int rc = check_instance(b, i);
  if (rc != GD_OK) return rc;
  if (!fields) return b->fail(GD_ERR_ARG, "null field table");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return b->fail(GD_ERR_ARG, "null field array");
  BH(b, hipStreamSynchronize(b->stream));
  const Inst& in = b->inst[i];
  for (int c = 0; c < 6; ++c)
    BH(b, hipMemcpy(fields[c], in.f[in.parity][c], in.N * sizeof(float), hipMemcpyDeviceToHost));
  return GD_OK;
}

int gd_batch_reset(gd_batch* b) {
  if (!b) return GD_ERR_ARG;
  BH(b, hipStreamSynchronize(b->stream));
  for (auto& in : b->inst) {
    for (int k = 0; k < 2; ++k)
      for (int c = 0; c < 6; ++c)
        BH(b, hipMemsetAsync(in.f[k][c], 0, in.N * sizeof(float), b->stream));
    in.parity = 0;
  }
  BH(b, hipMemsetAsync(b->d_step, 0, sizeof(int), b->stream));
  BH(b, hipMemsetAsync(b->d_parity, 0, gd_batch::kMaxGroups * sizeof(int), b->stream));
  BH(b, hipStreamSynchronize(b->stream));
  b->step = 0;
  return GD_OK;
}

int gd_batch_get_step(const gd_batch* b, size_t* step) {
  if (!b || !step) return GD_ERR_ARG;
  *step = b->step;
  return GD_OK;
}

int gd_batch_num_classes(const gd_batch* b, int* n) {
  if (!b || !n) return GD_ERR_ARG;
  *n = (int)b->cls.size();
  return GD_OK;
}

int gd_batch_class_info(const gd_batch* b, int c, size_t dims[3], int* ninst) {
  if (!b || c < 0 || c >= (int)b->cls.size()) return GD_ERR_ARG;
  if (dims) {
    dims[0] = b->cls[c].nx;
    dims[1] = b->cls[c].ny;
    dims[2] = b->cls[c].nz;
  }
  if (ninst) *ninst = (int)b->cls[c].insts.size();
  return GD_OK;
}

int gd_batch_set_class_tiling(gd_batch* b, int c, const gd_tiling* t) {
  if (!b) return GD_ERR_ARG;
  if (c < 0 || c >= (int)b->cls.size() || !t) return b->fail(GD_ERR_ARG, "bad class or tiling");
  const gd::ZTiling z{t->blt, t->rows, t->waves, t->zchunk};
  if (!gd::ztiling_supported(z, gd::kF32, gd::kMatIso))
    return b->fail(GD_ERR_UNSUPPORTED, "unsupported tiling");
  invalidate(b);
  b->cls[c].t = *t;
  b->cls[c].tuned = false;
  return GD_OK;
}

int gd_batch_get_class_tiling(const gd_batch* b, int c, gd_tiling* t) {
  if (!b || c < 0 || c >= (int)b->cls.size() || !t) return GD_ERR_ARG;
  *t = b->cls[c].t;
  return GD_OK;
}

int gd_batch_autotune(gd_batch* b, int reps) {
  if (!b) return GD_ERR_ARG;
  if (reps < 1) return b->fail(GD_ERR_ARG, "reps must be >= 1");
  for (auto& in : b->inst)
    if (!in.have_mat) return b->fail(GD_ERR_STATE, "every instance needs a material");
  // timing launches advance levels from the current step: the source tables
  // must cover blt more steps
  size_t cover = (size_t)-1;
  for (auto& in : b->inst)
    for (auto& s : in.src) cover = std::min(cover, s.len > b->step ? s.len - b->step : 0);
  if (cover == 0) return b->fail(GD_ERR_ARG, "source tables end at the current step");
  int rc = upload_sources(b);
  if (rc == GD_OK) rc = normalize_parity(b);
  if (rc != GD_OK) return rc;
  invalidate(b);
  unsigned* d_cnt = nullptr;
  const size_t ncnt = (size_t)kCountersPerLaunch * (reps + 1);
  BH(b, hipMalloc(&d_cnt, ncnt * sizeof(unsigned)));
  const int rw[4][2] = {{2, 4}, {2, 8}, {4, 4}, {4, 8}};
  for (auto& sc : b->cls) {
    std::vector<int> insts = sc.insts;
    double cells = 0.0;
    for (int i : insts) cells += (double)b->inst[i].N;
    std::vector<int> zcs = {16, 32, 64, (int)sc.nz};
    std::sort(zcs.begin(), zcs.end());
    zcs.erase(std::unique(zcs.begin(), zcs.end()), zcs.end());
    double best = 1e300;
    gd_tiling best_t = sc.t;
    std::vector<std::pair<double, gd_tiling>> cand;  // time per cell-update, tiling
    for (int blt = 1; blt <= 4; ++blt) {
      if ((size_t)blt > cover) break;
      for (auto& q : rw)
        for (int zc : zcs) {
          if (zc > (int)sc.nz) continue;
          gd::ZTiling t{blt, q[0], q[1], zc};
          if (!gd::ztiling_supported(t, gd::kF32, gd::kMatIso)) continue;
          Plan p;
          rc = build_plan(b, insts, t, false, p);
          if (rc != GD_OK) {
            free_plan(p);
            (void)hipFree(d_cnt);
            return rc;
          }
          hipError_t e = hipMemsetAsync(d_cnt, 0, ncnt * sizeof(unsigned), b->stream);
          size_t nl = 0;
          for (int r = 0; r <= reps && e == hipSuccess; ++r) {
            if (r == 1) e = hipEventRecord(b->ev0, b->stream);
            gd::BatchLaunch L{};
            L.step = b->d_step;
            L.parity = b->d_parity;  // 0: read copy 0, write copy 1
            L.step_offset = 0;
            L.launch_index = 0;
            if (e == hipSuccess) {
              rc = launch_plan(b, p, L, d_cnt + kCountersPerLaunch * r, b->stream, &nl);
              if (rc != GD_OK) {
                free_plan(p);
                (void)hipFree(d_cnt);
                return rc;
              }
            }
          }
          if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
          if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
          float ms = 0.0f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
          free_plan(p);
          if (e != hipSuccess) {
            (void)hipFree(d_cnt);
            return b->hip_fail(e, "autotune launch");
          }
          const double per_cell = (double)ms / reps / (cells * blt);
          cand.emplace_back(per_cell, gd_tiling{blt, q[0], q[1], zc});
          if (per_cell < best) {
            best = per_cell;
            best_t = gd_tiling{blt, q[0], q[1], zc};
          }
        }
    }
    // Within 3 % of the fastest, the deepest temporal blocking, then the
    // tallest tile, then the fastest: a class is timed alone, while in a
    // batch run the launch groups run at once and share HBM, where the fewer
    // bytes of more levels per launch and of smaller tile rims count more
    // (256-instance batch: 64^3 at B = 3, the fastest alone, 215.5 Gcell/s,
    // at B = 4 218.2-219.9; 160^3 at 2 x 8 instead of 4 x 8 208-212).
    int deepest = 0, tallest = 0;
    double pick_t = 1e300;
    for (const auto& c : cand) {
      const int rows = c.second.rows * c.second.waves;
      if (c.first <= best * 1.03 &&
          (c.second.blt > deepest || (c.second.blt == deepest && rows > tallest) ||
           (c.second.blt == deepest && rows == tallest && c.first < pick_t))) {
        deepest = c.second.blt;
        tallest = rows;
        pick_t = c.first;
        best_t = c.second;
      }
    }
    sc.t = best_t;
    sc.tuned = true;
  }
  (void)hipFree(d_cnt);
  return GD_OK;
}

int gd_batch_run(gd_batch* b, size_t nsteps, int mode) {
  if (!b) return GD_ERR_ARG;
  if (mode != GD_BATCH_PERSISTENT && mode != GD_BATCH_GRAPH && mode != GD_BATCH_PER_INSTANCE)
    return b->fail(GD_ERR_ARG, "unknown batch run mode");
  for (auto& in : b->inst) {
    if (!in.have_mat) return b->fail(GD_ERR_STATE, "every instance needs a material");
    for (auto& s : in.src)
      if (b->step + nsteps > s.len)
        return b->fail(GD_ERR_ARG, "source table does not cover the requested steps");
  }
  if (b->step + nsteps > (size_t)INT32_MAX) return b->fail(GD_ERR_ARG, "step count too large");
  b->last_ms = 0.0;
  b->last_nlaunch = 0;
  if (nsteps == 0) return GD_OK;
  int rc = ensure_plan(b);
  if (rc != GD_OK) return rc;
  rc = ensure_streams(b, b->groups.size());
  if (rc != GD_OK) return rc;
  Schedule* s = nullptr;
  rc = get_schedule(b, nsteps, &s);
  if (rc != GD_OK) return rc;

  size_t nl = 0;
  BH(b, hipEventRecord(b->ev0, b->stream));
  if (mode == GD_BATCH_PER_INSTANCE) {
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
      Group& g = b->groups[gi];
      for (int i : g.insts) {
        const Inst& in = b->inst[i];
        int par = in.parity;
        for (int k = 0; k < s->nlaunch[gi]; ++k) {
          const int blt = (int)std::min<size_t>((size_t)g.blt, nsteps - (size_t)k * g.blt);
          gd::ZTiling t{blt, g.rows, g.waves, b->cls[in.cls].t.zchunk};
          gd::ZArgs a = make_args(b, i, par, t);
          BH(b, gd::launch_zstream(t, gd::kF32, a.src, a.dst, a.sc, a.m, batch_mat(b), a.g, a.s,
                                   a.d, (int)(b->step + (size_t)k * g.blt), 0,
                                   gd::ZPart{gd::kZFull, nullptr, 0}, b->stream));
          par ^= 1;
          ++nl;
        }
      }
    }
    // keep the device step and group parities in step with the host
    BH(b, gd::launch_batch_advance(b->d_step, (int)nsteps, b->d_parity, s->d_flip,
                                   (int)b->groups.size(), s->d_counters, 0, b->stream));
    ++nl;
  } else if (mode == GD_BATCH_PERSISTENT) {
    rc = issue_run(b, nsteps, *s, &nl);
    if (rc != GD_OK) return rc;
  } else {
    if (!s->exec) {
      hipGraph_t graph = nullptr;
      BH(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      // the capture starts after ev0 was recorded on the stream
      rc = issue_run(b, nsteps, *s, &nl);
      hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (rc != GD_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return b->hip_fail(e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&s->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return b->hip_fail(e, "hipGraphInstantiate");
    }
    BH(b, hipGraphLaunch(s->exec, b->stream));
    nl = 0;
    for (int L : s->nlaunch) nl += (size_t)L;
    nl += 1;
  }
  BH(b, hipEventRecord(b->ev1, b->stream));
  BH(b, hipEventSynchronize(b->ev1));
  float ms = 0.0f;
  BH(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->last_nlaunch = nl;
  for (size_t gi = 0; gi < b->groups.size(); ++gi)
    if (s->nlaunch[gi] & 1)
      for (int i : b->groups[gi].insts) b->inst[i].parity ^= 1;
  b->step += nsteps;
  return GD_OK;
}

int gd_batch_last_run_stats(const gd_batch* b, double* total_ms, size_t* nlaunch) {
  if (!b) return GD_ERR_ARG;
  if (total_ms) *total_ms = b->last_ms;
  if (nlaunch) *nlaunch = b->last_nlaunch;
  return GD_OK;
}

int gd_batch_device_bytes(const gd_batch* b, size_t* bytes) {
  if (!b || !bytes) return GD_ERR_ARG;
  size_t t = b->arena_bytes;
  for (auto& in : b->inst)
    for (auto& s : in.src) t += s.len * sizeof(float);
  t += b->inst.size() * gd::kMaxSources * sizeof(gd::SourceDesc);
  *bytes = t;
  return GD_OK;
}

}  // extern "C"
