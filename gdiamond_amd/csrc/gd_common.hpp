// Structures shared by the host dispatcher and the gfx950 stepping kernels.
//
// Arithmetic contract (all kernels): every operation rounded once (the
// library is compiled with -ffp-contract=off), operand order of the reference
// update (gdiamond.hpp:1193-1221):
//   E = Ca*E + Cb*(((Hp - Hq) - (Hr - Hs)) - J*dx)
//   H = Da*H + Db*(((Ep - Eq) - (Er - Es)) - M*dx)
// computed in the storage format's compute type (FP32 for FP32, FP16 and
// BF16 storage, FP64 for FP64 storage), so results are bit-identical to the
// sequential CPU update of the same storage format.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

namespace gd {

constexpr int kMaxSources = 8;
constexpr int kMaxBlt = 8;

// Storage formats of the field copies (and of the isotropic coefficient).
//   kF32: float, computed in float (the reference's format).
//   kF64: double, computed in double (reference runs for error studies).
//   kF16: IEEE binary16 of value * 2^-k_c per component c; values are
//         decoded exactly to float, computed in float and rounded to binary16
//         (round to nearest even) when stored: after every half step in the
//         PLAIN kernels, once per launch of B steps in the TILED kernel.
//   kBF16: bfloat16 (the upper half of a float) of value * 2^-k_c, same
//         decode / compute / round-when-stored rules as kF16.
enum Storage { kF32 = 0, kF64 = 1, kF16 = 2, kBF16 = 3 };

// Field arrays of one field copy (SoA, idx = i + j*nx + k*nx*ny); the element
// type follows the storage format.
struct Fields {
  void* c[6];  // Ex Ey Ez Hx Hy Hz
};

// value = stored * s[c] for 16-bit storage (s = 2^k_c, exact), inv = 1/s.
struct Scales {
  float s[6];
  float inv[6];
};

// Material kinds.
//   kMatIso:  isotropic lossless, Ca = Da = 1 (multiplication by 1 is exact and
//             omitted), cb per cell in the compute type (FP32; FP64 for
//             FP64 storage), db scalar.
//   kMatFull: the reference's 12 arrays Cax Cbx Cay Cby Caz Cbz Dax Dbx Day Dby
//             Daz Dbz in the compute type.
//   kMatIdx:  isotropic lossless with a 1-byte material id per cell and an FP32
//             table of at most 256 Cb values (exact Cb, 1 byte per cell).
enum MatKind { kMatIso = 0, kMatFull = 1, kMatIdx = 2 };
constexpr int kMaxMaterials = 256;

// Host vector whose elements start uninitialised: large staging arrays are
// then first touched by the threads that fill them, not zeroed on one thread.
template <class T>
struct DefaultInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = DefaultInitAlloc<U>;
  };
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    if constexpr (sizeof...(A) == 0)
      ::new ((void*)p) U;
    else
      ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
template <class T>
using HostVec = std::vector<T, DefaultInitAlloc<T>>;

// Ids of float bit patterns in order of first insertion, at most
// kMaxMaterials of them (open addressing over 1024 slots, load <= 1/4).
struct BitsIndex {
  static constexpr unsigned kSlots = 1024;
  uint32_t key[kSlots];
  int val[kSlots];
  std::vector<uint32_t> order;  // bit patterns by id
  BitsIndex() {
    for (unsigned q = 0; q < kSlots; ++q) val[q] = -1;
  }
  // id of `bits`, inserted as the next id when new; -1 when new and all
  // kMaxMaterials ids are taken
  int find_or_insert(uint32_t bits) {
    unsigned h = (bits * 2654435761u) >> 22;
    while (val[h] >= 0 && key[h] != bits) h = (h + 1) & (kSlots - 1);
    if (val[h] < 0) {
      if ((int)order.size() == kMaxMaterials) return -1;
      key[h] = bits;
      val[h] = (int)order.size();
      order.push_back(bits);
    }
    return val[h];
  }
};

// Material ids of an isotropic Cb array when it holds at most
// kMaxMaterials distinct bit patterns (ids in order of first appearance).
// Returns false when there are more.
// Up to 16 host threads each number the values of one contiguous range in
// their own first-appearance order; merging the ranges' lists in range order
// then gives every value its global first-appearance id (a value first
// appears in the earliest range that holds it, and within that range in its
// local order), and the ranges whose local ids differ are renumbered.
inline bool index_materials(const float* cb, size_t n, HostVec<uint8_t>& id,
                            std::vector<float>& table) {
  id.resize(n);
  table.clear();
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, n >> 20));
  const size_t chunk = (n + nt - 1) / nt;
  std::vector<BitsIndex> loc(nt);
  std::vector<char> over(nt, 0);
  auto run = [&](auto&& f) {
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) th.emplace_back([&f, t] { f(t); });
    f(0u);
    for (auto& x : th) x.join();
  };
  run([&](unsigned t) {
    const size_t b = std::min(n, (size_t)t * chunk), e = std::min(n, b + chunk);
    uint32_t last_bits = 0;
    int last_id = -1;
    for (size_t q = b; q < e; ++q) {
      uint32_t bits;
      std::memcpy(&bits, &cb[q], 4);
      if (last_id < 0 || bits != last_bits) {
        last_id = loc[t].find_or_insert(bits);
        if (last_id < 0) {
          over[t] = 1;
          return;
        }
        last_bits = bits;
      }
      id[q] = (uint8_t)last_id;
    }
  });
  for (unsigned t = 0; t < nt; ++t)
    if (over[t]) return false;
  BitsIndex glob;
  std::vector<std::vector<uint8_t>> map(nt);
  std::vector<char> ident(nt, 1);
  for (unsigned t = 0; t < nt; ++t) {
    for (size_t k = 0; k < loc[t].order.size(); ++k) {
      const int g = glob.find_or_insert(loc[t].order[k]);
      if (g < 0) return false;
      map[t].push_back((uint8_t)g);
      if ((size_t)g != k) ident[t] = 0;
    }
  }
  run([&](unsigned t) {
    if (ident[t]) return;
    const size_t b = std::min(n, (size_t)t * chunk), e = std::min(n, b + chunk);
    const uint8_t* m = map[t].data();
    for (size_t q = b; q < e; ++q) id[q] = m[id[q]];
  });
  table.resize(glob.order.size());
  std::memcpy(table.data(), glob.order.data(), table.size() * sizeof(float));
  return true;
}

struct Material {
  const void* c[12];
  const void* cb;
  const uint8_t* id;     // kMatIdx: material id per cell
  const float* table;    // kMatIdx: Cb of each id (device, kMaxMaterials entries)
  float db;
  double db64;
};

// Decode / encode between a storage element and its compute type.
template <class S>
struct Codec;
template <>
struct Codec<float> {
  using T = float;
  __device__ __forceinline__ static T dec(float v, float) { return v; }
  __device__ __forceinline__ static float enc(T v, float) { return v; }
  // value after one store-load round trip
  __device__ __forceinline__ static T round(T v, float, float) { return v; }
};
template <>
struct Codec<double> {
  using T = double;
  __device__ __forceinline__ static T dec(double v, float) { return v; }
  __device__ __forceinline__ static double enc(T v, float) { return v; }
  __device__ __forceinline__ static T round(T v, float, float) { return v; }
};
// binary16 storage of value * 2^-k.
template <>
struct Codec<_Float16> {
  using T = float;
  // Both directions are one mixed-precision FMA with a -0 addend: the binary16
  // operand converts to float exactly, the product by a power of two is exact
  // for the scales the library accepts (|exponent| <= 100 keeps it normal and
  // finite), and x + (-0) == x bit for bit, signed zeros included. So dec is
  // (float)v * s and enc is (_Float16)(v * inv) rounded once to nearest even,
  // in one instruction each instead of a convert and a multiply. (A +0
  // addend, which the compiler picks when it fuses the pair by itself, would
  // turn -0 into +0.)
  __device__ __forceinline__ static T dec(_Float16 v, float s) {
    float r;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(v), "s"(s), "v"(-0.0f));
    return r;
  }
  __device__ __forceinline__ static _Float16 enc(T v, float inv) {
    _Float16 r;
    asm("v_fma_mixlo_f16 %0, %1, %2, %3" : "=v"(r) : "v"(v), "s"(inv), "v"(-0.0f));
    return r;
  }
  __device__ __forceinline__ static T round(T v, float s, float inv) {
    return dec(enc(v, inv), s);
  }
};
// bfloat16 storage of value * 2^-k: decode widens the 16 stored bits (the
// upper half of a float) exactly and multiplies by 2^k in float; encode
// multiplies by 2^-k in float and rounds to the nearest bfloat16, ties to even
// (v_cvt_pk_bf16_f32). bfloat16 has float's exponent range, so the scale only
// matters for values near the ends of that range.
template <>
struct Codec<__bf16> {
  using T = float;
  __device__ __forceinline__ static T dec(__bf16 v, float s) {
    const uint32_t u = (uint32_t)__builtin_bit_cast(uint16_t, v) << 16;
    return __builtin_bit_cast(float, u) * s;
  }
  __device__ __forceinline__ static __bf16 enc(T v, float inv) {
    return (__bf16)(v * inv);
  }
  __device__ __forceinline__ static T round(T v, float s, float inv) {
    return dec(enc(v, inv), s);
  }
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
  float* snap;       // TILED: plane values per level, [level][slot][plane]
  int snap_base;     // TILED: level slot of the launch's first level
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
__device__ __forceinline__ void dft_accumulate(const Dft& d, int slot, long long p, double dv,
                                               int run_step) {
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
__device__ __forceinline__ void dft_accumulate_group(const Dft& d, long long p, double v0,
                                                     double v1, double v2, int run_step) {
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

// ---- device helpers shared by the stepping kernels

// Lane l receives the value of lane l-1 (lane 0 receives 0): wave_shr:1.
// bound_ctrl writes 0 for the lane without a source, which lets the compiler
// fold the shift into the subtraction that consumes it (v_sub*_f32_dpp).
__device__ __forceinline__ float from_lane_below(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
// Lane l receives the value of lane l+1 (lane 63 receives 0): wave_shl:1.
__device__ __forceinline__ float from_lane_above(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, true));
}
// The same shifts for doubles, one DPP move per 32-bit half.
template <int kCtrl>
__device__ __forceinline__ double shift_lanes_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, kCtrl, 0xF, 0xF,
                                                            true);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), kCtrl,
                                                            0xF, 0xF, true);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double from_lane_below(double v) { return shift_lanes_f64<0x138>(v); }
__device__ __forceinline__ double from_lane_above(double v) { return shift_lanes_f64<0x130>(v); }

// Workgroups are dispatched round-robin over the 8 XCDs. Give each XCD a
// contiguous run of tiles (x fastest, then y, then z) so tiles that share
// halos share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n >> 3, r = n & 7;
  const int x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Load / store at a per-plane base (uniform, SGPRs) plus a 32-bit byte offset
// (per lane), so the hardware's scalar-base + vector-offset addressing is used
// instead of 64-bit per-lane address arithmetic. The offsets are computed
// once before the stream loop; without the empty asm the compiler carries
// them into the loop as 64-bit values and no longer sees that the high half
// is zero, so every 4- or 8-byte access in the loop would cost a 64-bit add.
// (For 2-byte elements the plain form measures faster, and so it does for
// every element size in the FUSED kernel, which passes kScalarBase = false.)
//
// Every field, material and snapshot array lives in global device memory, and
// the accesses say so (gptr): a pointer the compiler cannot prove global (one
// read from memory, as the batch kernel reads its per-instance arguments)
// would otherwise be accessed with flat instructions, which also count in the
// LDS wait counter, so that each wait for an LDS read (the exchanged rows,
// the Cb table) waited for the prefetched plane as well (256-instance batch
// with flat accesses 209-213, with global ones 215-219 Gcell/s).
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}
template <class S, bool kScalarBase = (sizeof(S) >= 4)>
__device__ __forceinline__ S ldp(const void* base, size_t plane_off, unsigned off) {
  if constexpr (kScalarBase) asm("" : "+v"(off));
  return *gptr(reinterpret_cast<const S*>(static_cast<const char*>(base) + plane_off + off));
}
template <class S, bool kScalarBase = (sizeof(S) >= 4)>
__device__ __forceinline__ void stp(void* base, size_t plane_off, unsigned off, S v) {
  if constexpr (kScalarBase) asm("" : "+v"(off));
  *gptr(reinterpret_cast<S*>(static_cast<char*>(base) + plane_off + off)) = v;
}

// Store through a buffer resource spanning one plane (`plane_bytes` < 2^31
// bytes from base + plane_off). A lane whose offset has bit 31 set
// (kStoreDrop) lies past the resource's range, and the hardware drops its
// store. Every lane of the wave then issues the same store instruction, so a
// loop that stores only some cells on some planes still issues a fixed number
// of vector-memory instructions per iteration. The compiler can then wait for
// loads issued before those stores without also waiting for the stores
// (the vector-memory counter completes in issue order).
// Used for 4-byte elements only (16-bit fields measured slower this way).
constexpr unsigned kStoreDrop = 0x80000000u;
template <class S>
__device__ __forceinline__ void stp_drop(void* base, size_t plane_off, unsigned plane_bytes,
                                         unsigned off, S v) {
  static_assert(sizeof(S) == 4, "4-byte elements");
  // dword 3 of the gfx9 buffer descriptor: 32-bit data format, no swizzle
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      static_cast<char*>(base) + plane_off, (short)0, (int)plane_bytes, 0x00020000);
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, (int)off, 0, 0);
}

}  // namespace gd

struct gd_nn;
// Host-side launchers (defined in the .hip kernel files).
namespace gd {
hipError_t launch_plain_step(Storage st, const Fields& f, const Scales& sc, const Material& m,
                             MatKind mk, const Grid& g, const Sources& s, const Dft& d, int n,
                             int run_step, hipStream_t stream);
// One step of the FUSED kernel: fields of step n from `in`, step n+1 to `out`.
hipError_t launch_fused_step(Storage st, const Fields& in, const Fields& out, const Scales& sc,
                             const Material& m, MatKind mk, const Grid& g, const Sources& s,
                             const Dft& d, int n, int run_step, hipStream_t stream);

// zsplit >= 0: the z chunk holding plane zsplit is cut into the part below
// it, the one-plane chunk [zsplit, zsplit + 1) and the part above it (see
// chunk_range); -1: uniform chunks of zchunk planes.
struct ZTiling {
  int blt, rows, waves, zchunk;
  int zsplit = -1;
};

// Planes [lo, hi) of z chunk tz: chunks of zc planes, except that with
// zsplit >= 0 the chunk holding plane zsplit is cut into [start, zsplit)
// (when not empty), [zsplit, zsplit + 1) and [zsplit + 1, end) (when not
// empty). A TILED launch with a monitored z plane cuts there, so that only
// the tiles of the one-plane chunk run the build with the snapshot code, and
// every other tile of that layer runs the lean build.
__host__ __device__ inline void chunk_range(int zc, int nz, int zsplit, int tz, int& lo, int& hi) {
  if (zsplit < 0) {
    lo = tz * zc;
    hi = lo + zc < nz ? lo + zc : nz;
    return;
  }
  const int c = zsplit / zc;
  const int end = (c + 1) * zc < nz ? (c + 1) * zc : nz;
  const int below = zsplit > c * zc ? 1 : 0;
  const int above = zsplit + 1 < end ? 1 : 0;
  int k = tz - c;
  if (k < 0) {
    lo = tz * zc;
    hi = lo + zc;
    return;
  }
  if (below) {
    if (k == 0) {
      lo = c * zc;
      hi = zsplit;
      return;
    }
    --k;
  }
  if (k == 0) {
    lo = zsplit;
    hi = zsplit + 1;
    return;
  }
  --k;
  if (above) {
    if (k == 0) {
      lo = zsplit + 1;
      hi = end;
      return;
    }
    --k;
  }
  lo = (c + 1 + k) * zc;
  hi = lo + zc < nz ? lo + zc : nz;
}
// Number of z chunks of chunk_range.
__host__ __device__ inline int chunk_count(int zc, int nz, int zsplit) {
  const int n = (nz + zc - 1) / zc;
  if (zsplit < 0) return n;
  const int c = zsplit / zc;
  const int end = (c + 1) * zc < nz ? (c + 1) * zc : nz;
  return n + (zsplit > c * zc ? 1 : 0) + (zsplit + 1 < end ? 1 : 0);
}

// Arguments of one TILED launch over one problem: source and destination
// field copies, decomposition into ntx*nty*ntz tiles of zc planes.
struct ZArgs {
  Fields src;
  Fields dst;
  Scales sc;
  Material m;
  Grid g;
  Sources s;
  Dft d;
  int n0;         // absolute step of level 1
  int run_step0;  // DFT weight column of level 1
  int ntx, nty, ntz;
  int zc;         // planes owned per workgroup
  int scw;        // tile order: 0 = x fastest along whole rows; w > 0 = columns
                  // of w tiles, x fastest within a column, when ntx > w
  int zsplit;     // chunk_range's cut plane, -1 for uniform chunks
};

// Tile decomposition of an nx*ny*nz problem under a tiling.
inline void ztile_counts(const ZTiling& t, int nx, int ny, int nz, int& ntx, int& nty, int& ntz) {
  const int ox = 64 - 2 * t.blt;
  const int oy = t.rows * t.waves - 2 * t.blt;
  ntx = (nx + ox - 1) / ox;
  nty = (ny + oy - 1) / oy;
  ntz = chunk_count(t.zchunk, nz, t.zsplit);
}

// Batched persistent launch: every workgroup takes (instance, tile) work items
// from a queue until it is empty. Instance descriptors come in pairs (field
// copy parity 0 and 1); the absolute step and each group's parity live in
// device memory so that a captured graph can be replayed unchanged.
struct BatchItem {
  int inst;
  int tile;
  int full;  // 1: the tile holds a source cell (runs the build with source code)
};
struct BatchLaunch {
  const ZArgs* inst;     // [2 * ninst]: src = copy p, dst = copy 1-p for parity p
  const BatchItem* items;
  int nitems;
  unsigned* counter;     // queue position of this launch; zero on entry
  const int* step;       // absolute step at the start of the run
  const int* parity;     // copy parity of the group at the start of the run
  int step_offset;       // steps of the run before this launch
  int launch_index;      // launches of the group before this one in the run
};
// One persistent launch over the items of L; each item runs the tile with or
// without the point-source code as its `full` flag says.
hipError_t launch_zstream_batch(const ZTiling& t, MatKind mk, const BatchLaunch& L, int nblocks,
                                hipStream_t stream);
// Workgroups of one batch launch that fill the device (occupancy x CUs).
int zstream_batch_blocks(const ZTiling& t, MatKind mk, int device);
// End of a run: step += nsteps, parity[g] ^= flip[g], every queue counter = 0.
hipError_t launch_batch_advance(int* step, int nsteps, int* parity, const int* flip, int ngroups,
                                unsigned* counters, int ncounters, hipStream_t stream);
bool ztiling_supported(const ZTiling& t, Storage st, MatKind mk);
// Kernel forms of a TILED launch: kZFull runs tiles with the point-source and
// DFT code, kZLean without it (for tiles that need neither). `list` (device,
// nlist tile ids tx + ntx*(ty + nty*tz)) restricts the launch to those
// tiles; null runs every tile. The two lists of one classification
// (zstream_classify) together cover every tile once.
enum ZForm { kZFull = 0, kZLean = 1 };
// One workgroup of the lean form that runs several smaller tiles of one z
// chunk side by side: lane groups of g lanes (64, 32 or 16) times h wave
// groups (1 or 2) of W/h waves; sub-tile (q, k) is tile (tx[q][k], ty[q][k])
// and owns cells when bit 2q+k of `own` is set. A lean list entry -1-i names
// packs[i].
struct ZPack {
  int tz;
  int g;
  int h;
  int own;
  int tx[4][2];
  int ty[4][2];
};
struct ZPart {
  int form;
  const int* list;
  int nlist;
  const ZPack* packs = nullptr;
};
// The lean form exists for the 4 x 8 tile (the planner's).
bool zstream_lean_supported(const ZTiling& t);
// Column width of the tile order (ZArgs::scw) of single-problem launches for
// a storage format; batch launches use 0 (x fastest along whole rows).
int zstream_scw(Storage st);
// Host: splits the tiles of tiling t into those that need the source or DFT
// code (a source cell in the tile's input window and streamed z range, or
// owned cells on the monitored plane) and the rest, each list of tile ids in
// the tile order of column width scw.
void zstream_classify(const ZTiling& t, int scw, const Grid& g, const SourceDesc* src, int nsrc,
                      const Dft& d, std::vector<int>& extra, std::vector<int>& lean);
// Host: the work items of one single-problem launch pair. `extra`: ids of the
// tiles that need the source or DFT code (full form), in the kernels' tile
// order. `lean`: tile ids, or -1-i for packs[i], which gathers tiles of the
// remainder column or row (when the tile size leaves a narrow or short one)
// into shared workgroups, in chunk-group order (zstream_chunk_groups) for
// FP32 fields and in the tile order of zstream_scw(st) otherwise.
void zstream_plan_items(const ZTiling& t, Storage st, const Grid& g, const SourceDesc* src,
                        int nsrc, const Dft& d, std::vector<int>& extra, std::vector<int>& lean,
                        std::vector<ZPack>& packs);
// Host: reorders a lean list (items of ntz z chunks of ncol x-y columns each,
// chunk by chunk in the tile order) into groups of gc consecutive columns,
// each group walked through every z chunk before the next group. Each XCD
// takes a contiguous run of the list and holds 32 workgroups at a time, so a
// chunk starts gc items, i.e. gc/32 of a chunk's stream, after the chunk
// below it in the same columns; with gc = 32 D / (D + 2B) (D planes per
// chunk, B levels) that is when the chunk below reaches the 2B planes the two
// share (its top halo, the other's bottom halo), so both read those planes
// at about the same time. (FP32, 32-plane chunks: 512^3 227 -> 236
// Gcell/s, 768^3 236 -> 244, 1024^3 233 -> 243 and 234 -> 253 on two
// boxes. The L2 fetch count per launch does not drop, 6.4 GB before and
// after at 512^3, so the second of the two simultaneous reads is served
// below L2; the timing matters: 20 or 32 columns per group instead of 27 at
// 48-plane chunks ran 4-5 % slower than the plain order.)
void zstream_chunk_groups(const ZTiling& t, int ntz, int ncol, std::vector<int>& lean,
                          const std::vector<ZPack>& packs);
hipError_t launch_zstream(const ZTiling& t, Storage st, const Fields& src, const Fields& dst,
                          const Scales& sc, const Material& m, MatKind mk, const Grid& g,
                          const Sources& s, const Dft& d, int n0, int run_step0,
                          const ZPart& part, hipStream_t stream);

hipError_t launch_copy(const float* src, float* dst, size_t n, hipStream_t stream);

// Learned corrector (gd_learned.hip). Feature and output scales of the
// network: features E*se, H*sh, (E-E_prev)*sd; corrections y/re, y/rh.
struct NnFeat {
  float se, sh, sd;
  float re, rh;
  float dx_over_dt;  // 1/eps = cb * dx/dt
};
hipError_t launch_restrict(const Fields& f, int nx, int ny, int nz, float* out, hipStream_t s);
// out[q] = table[id[q]] for q < n: the dense Cb of an indexed material.
hipError_t launch_expand_material(const uint8_t* id, const float* table, size_t n, float* out,
                                  hipStream_t s);
int nn_dims(const ::gd_nn* o, int* nx, int* ny, int* nz);
// features -> network -> corrected fields, all on stream s
hipError_t nn_correct(::gd_nn* o, const Fields& f, const float* pe0, const float* pe1,
                      const float* pe2, const float* cb, const NnFeat& nf, hipStream_t s);

// Accumulates the DFT of `nlev` consecutive steps from the per-level plane
// snapshot the TILED kernel wrote (run steps run_step0 .. run_step0+nlev-1).
hipError_t launch_dft_snapshot(const Dft& d, int nlev, int run_step0, hipStream_t stream);
}  // namespace gd
