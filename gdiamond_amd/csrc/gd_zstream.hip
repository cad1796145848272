// TILED stepping kernel for gfx950: 2.5-D temporal blocking.
//
// Replaces the reference's mix-mapping tilings (ver2 diamond,
// kernels_mm_ver2.cuh:488-720; ver3 overlapped "replication" mountains,
// kernels_mm_ver3.cuh:417-647). Those load one 16x8x8 block of all six fields
// into shared memory and run BLT sub-steps on it. Here a workgroup owns an
// overlapped x-y tile and streams it along z through BLT time levels kept in
// registers, so each field value is read from HBM once per launch and written
// once, whatever BLT is, and redundant compute exists only in x and y.
//
// Geometry of one workgroup (template B = time levels, R = rows per thread,
// W = waves):
//   * input tile 64 x (W*R) columns; lane = x (one wave64 per x row segment),
//     each thread holds R consecutive y rows of its lane in registers;
//   * owned output 64-2B x W*R-2B columns (the halo shrinks by one cell per
//     side per level: E needs H at i-1/j-1, H needs E at i+1/j+1);
//   * z: owned planes [z_lo, z_hi); streaming starts B planes below z_lo.
// Level b (1..B) runs one plane behind level b-1: at stream position p it
// updates E of step n0+b-1 on plane p-b+1 and then H on plane p-b.
// Neighbours: x via DPP lane shifts across the wave (wave_shr/wave_shl), y
// within a thread's rows from registers and across waves through a small LDS
// exchange of one row, z from the previous plane kept in registers.
// Physical boundary cells (index 0 or n-1 on any axis) are frozen, exactly as
// the reference updates only 1..n-2 (gdiamond.hpp:1194-1221).
//
// Validity: a cell of level b is exact when its whole dependency cone lies in
// the loaded tile; owned cells of level B always do. Cells outside the cone
// (tile rim, planes below the chunk, columns outside the grid) hold garbage
// that only ever feeds other cells outside the cone, so loads are clamped into
// the grid instead of predicated.
//
// Storage formats: values are decoded when loaded and encoded when stored;
// the B levels in between stay in FP32 registers. For binary16 storage a
// launch therefore rounds to binary16 once per B steps (at its store), and the
// DFT snapshots hold the FP32 value of every level. Every workgroup that
// computes a cell at some level does so from the same decoded inputs with the
// same operations, so overlapping tiles agree bit for bit.
//
// Two builds of the tile: zstream_kernel with the point-source and DFT code,
// and zstream_lean_kernel without it. The host sorts the tiles of a launch
// (zstream_classify) and runs the few that need that code with the full
// kernel next to the lean kernel over the rest; the result is the same as one
// full launch, since the code it leaves out does nothing in those tiles.
#include <algorithm>
#include <type_traits>

#include "gd_common.hpp"

namespace gd {

namespace {

// Subtract the point-source terms of step n on plane z from the curl
// brackets c[0..2] (components base..base+2) of row gy (uniform) at column gx.
// The loop over sources is uniform; a lane applies a term only on
// its own cell (at most one source per cell and component, so the order of
// the subtractions does not matter). Called only by waves that hold a source
// column.
__device__ __forceinline__ void apply_sources_row(const Sources& s, int gx, int gy, int z, int n,
                                                  float dx, int base, float (&c)[3]) {
  for (int q = 0; q < s.n; ++q) {
    const SourceDesc sd = s.d[q];
    const int k = sd.comp - base;
    if (sd.k != z || sd.j != gy || k < 0 || k > 2) continue;
    const float v = *gptr(sd.table + n);
    const bool m = sd.i == gx;
    c[0] = (m && k == 0) ? c[0] - v * dx : c[0];
    c[1] = (m && k == 1) ? c[1] - v * dx : c[1];
    c[2] = (m && k == 2) ? c[2] - v * dx : c[2];
  }
}

// Occupancy target per SIMD: the register state is ~6 values per row per time
// level, so 2-row threads fit 4 waves per SIMD (<= 128 VGPRs) and 4-row
// threads 2 waves.
template <int R>
constexpr int zstream_waves_per_simd() {
  return R <= 2 ? 4 : 2;
}

// Tile coordinates of launch position `tile`: x fastest, then y, then z
// chunk, or in columns of a.scw tiles (see ZArgs).
__host__ __device__ __forceinline__ void tile_coords(const ZArgs& a, const int tile, int& tx,
                                                     int& ty, int& tz) {
  if (a.scw <= 0 || a.ntx <= a.scw) {
    tx = tile % a.ntx;
    const int tyz = tile / a.ntx;
    ty = tyz % a.nty;
    tz = tyz / a.nty;
  } else {
    // super-columns of scw tiles (the last one narrower), walked x fastest,
    // then y, then the next column
    const int per_z = a.ntx * a.nty;
    tz = tile / per_z;
    const int rem = tile - tz * per_z;
    const int col = rem / (a.scw * a.nty);
    const int within = rem - col * a.scw * a.nty;
    const int left = a.ntx - col * a.scw;
    const int width = a.scw < left ? a.scw : left;
    ty = within / width;
    tx = col * a.scw + (within - ty * width);
  }
}

// LDS of one tile: exchange rows of every level and the Cb table of the
// indexed material. One object per (levels, waves, material form), shared by
// both builds of the tile, so that a kernel holding both (the batch kernel)
// allocates it once.
template <int B, int W, bool kIdx>
struct ZLds {
  // h[level-1][parity of p][wave][Hz,Hx][lane]: last row of H of a level,
  // read by the next wave for the j-1 neighbour of the next level.
  // e[level-1][parity of plane][wave][Ez,Ex][lane]: first row of E of a
  // level, read by the previous wave one stream position later (j+1).
  float h[B][2][W][2][64];
  float e[B][2][W][2][64];
  float cb[kIdx ? kMaxMaterials : 1];  // Cb of each material id
};
template <int B, int W, bool kIdx>
__device__ __forceinline__ ZLds<B, W, kIdx>& zlds() {
  __shared__ ZLds<B, W, kIdx> lds;
  return lds;
}

// One workgroup's work: tile (tx, ty, tz) of the problem described by `a`,
// streamed through B time levels. kX = false compiles out the point sources
// and the DFT snapshots (tiles without a source cell in their window and
// without owned monitored cells): their state no longer competes for
// registers in the stream loop.
//
// Sub-tiles: a workgroup may also run several smaller tiles side by side
// (the remainder column or row of a grid that the tile size does not divide):
// lanes form groups of G (64, 32 or 16) lanes and waves groups of WH (W or
// W/2) waves, each group pair its own tile (tx, ty), owning the cells at
// least B lanes and B rows inside its group. Lane shifts and the exchanged
// rows cross group edges and bring another tile's values there, but those
// reach only cells outside every owned cone, as the grid clamps do. A lane
// with `active` false owns nothing (padding of a partly filled workgroup).
//
// kSingle: the tile of a single-problem launch (fixed-count stores for
// 4-byte fields, material loaded ahead of the fields) or, false, of the
// batch kernel (branch-guarded stores, material loaded after the fields),
// the faster form for each.
template <int B, int R, int W, int kMat, class S, bool kX, bool kSingle = true>
__device__ __forceinline__ void zstream_tile(const ZArgs& a, const int tx, const int ty,
                                             const int tz, const int G = 64, const int WH = W,
                                             const bool active = true) {
  using C = Codec<S>;
  constexpr bool kFull = kMat == kMatFull;
  constexpr bool kIdx = kMat == kMatIdx;
  constexpr unsigned ES = sizeof(S);
  constexpr unsigned kLogES = ES == 4 ? 2 : (ES == 2 ? 1 : 0);
  constexpr int OX = 64 - 2 * B;
  constexpr int OY = W * R - 2 * B;
  constexpr bool kDropStores = ES == 4 && kSingle;
  ZLds<B, W, kIdx>& lds = zlds<B, W, kIdx>();
  auto& lds_h = lds.h;
  auto& lds_e = lds.e;
  auto& lds_cb = lds.cb;

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;

  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const size_t plane_bytes = (size_t)a.g.sxy * ES;
  const Scales& sc = a.sc;
  const float dx = a.g.dx;

  const int ll = lane & (G - 1);  // lane within its sub-tile (G: 64, 32 or 16)
  const int wl = w & (WH - 1);    // wave within its sub-tile (WH: W or W/2)
  const int gx = tx * OX - B + ll;
  const int gy0 = ty * OY - B + wl * R;
  const bool intx = gx >= 1 && gx <= nx - 2;
  const bool ownx = active && ll >= B && ll < G - B && gx < nx;
  const unsigned cgx = (unsigned)clampi(gx, 0, nx - 1);
  bool upd_xy[R], own[R];
  unsigned off[R];   // in-plane byte offset, clamped into the grid
  bool any_src = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int gy = gy0 + r;
    upd_xy[r] = intx && gy >= 1 && gy <= ny - 2;
    const int ly = wl * R + r;
    own[r] = ownx && ly >= B && ly < WH * R - B && gy < ny;
    off[r] = (cgx + (unsigned)clampi(gy, 0, ny - 1) * (unsigned)nx) * ES;
    for (int q = 0; q < a.s.n; ++q) {
      const SourceDesc sd = a.s.d[q];
      any_src = any_src || (sd.i == gx && sd.j == gy);
    }
  }
  const unsigned offm = (cgx + (unsigned)clampi(gy0 - 1, 0, ny - 1) * (unsigned)nx) * ES;
  // store offsets: the cell's offset for owned cells, else past the plane
  unsigned soff[R];
#pragma unroll
  for (int r = 0; r < R; ++r) soff[r] = own[r] ? off[r] : kStoreDrop;
  // wave-uniform: does any lane of this wave hold a source column?
  const bool wave_src = kX && __builtin_amdgcn_ballot_w64(any_src) != 0;

  int z_lo, z_hi;
  chunk_range(a.zc, nz, a.zsplit, tz, z_lo, z_hi);
  const int p_begin = max(0, z_lo - B);
  const int p_end = z_hi - 1 + B;

  // DFT monitor: a cell (x,y,z) of an owned row is on the monitored plane when
  // dft_row[r] and (axis != 2 or z == index); its plane index is
  // dft_base[r] + z*dft_zmul (xy plane: x + y*nx; yz plane: y + z*ny; xz
  // plane: x + z*nx).
  const bool dft_on = a.d.ncomp > 0;
  const int dft_zmul = a.d.axis == 2 ? 0 : (a.d.axis == 0 ? ny : nx);
  bool own_any = false, dft_any = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int gy = gy0 + r;
    own_any = own_any || own[r];
    const bool on = a.d.axis == 2 || (a.d.axis == 0 ? gx == a.d.index : gy == a.d.index);
    dft_any = dft_any || (dft_on && own[r] && on);
  }
  // wave-uniform: does this wave own any cell of the monitored plane?
  const bool wave_dft = kX && __builtin_amdgcn_ballot_w64(dft_any) != 0;
  // per-row monitor predicate and plane index base, evaluated only by waves
  // that own monitored cells
  auto dft_row = [&](int r) {
    const int gy = gy0 + r;
    return own[r] && (a.d.axis == 2 || (a.d.axis == 0 ? gx == a.d.index : gy == a.d.index));
  };
  auto dft_base = [&](int r) {
    const int gy = gy0 + r;
    return a.d.axis == 2 ? gx + gy * nx : (a.d.axis == 0 ? gy : gx);
  };

  // Pipeline state (registers):
  //   eq[b] = E of level b on plane p-b   (b = 1..B, slot 0 unused)
  //   hq[b] = H of level b on plane p-b-1 (b = 0..B-1; level 0 = loaded H)
  //   cbq[b-1] = Cb on plane p-b+1, used by level b (dense isotropic material)
  //   idq: material ids of planes p..p-B+1, one byte per level (byte b-1 for
  //        level b), looked up in lds_cb per level (indexed material)
  float eq[B + 1][R][3];
  float hq[B][R][3];
  float cbq[kIdx ? 1 : B][R];
  unsigned idq[R];
  static_assert(B <= 4, "material ids of B levels are packed in 32 bits");
#pragma unroll
  for (int b = 0; b <= B; ++b)
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        eq[b][r][c] = 0.0f;
        if (b < B) hq[b][r][c] = 0.0f;
      }
#pragma unroll
  for (int b = 0; b < (kIdx ? 1 : B); ++b)
#pragma unroll
    for (int r = 0; r < R; ++r) cbq[b][r] = 0.0f;
#pragma unroll
  for (int r = 0; r < R; ++r) idq[r] = 0u;

  // Exchange rows start at zero; they are only consumed by cells outside the
  // valid cone before the first real write, but stay deterministic.
  for (int q = threadIdx.x; q < B * 2 * W * 2 * 64; q += 64 * W) {
    (&lds_h[0][0][0][0][0])[q] = 0.0f;
    (&lds_e[0][0][0][0][0])[q] = 0.0f;
  }
  if constexpr (kIdx) {
    for (int q = threadIdx.x; q < kMaxMaterials; q += 64 * W) lds_cb[q] = *gptr(a.m.table + q);
  }

  // Level-0 values of the next stream position, loaded one iteration ahead so
  // their HBM latency overlaps the level updates of the current position
  // (fields kept in the storage format and decoded when consumed; the dense
  // Cb is FP32 whatever the field format). Planes >= nz are clamped to nz-1:
  // they only feed cells outside the cone.
  using MT = typename std::conditional<kIdx, uint8_t, float>::type;
  constexpr unsigned kCbMul = 4 / ES;  // FP32 Cb offsets from field offsets
  S pe[R][3], ph[R][3], phmz, phmx;
  MT pcb[R];
  auto prefetch_material = [&](size_t pz) {
    if constexpr (kIdx) {
#pragma unroll
      for (int r = 0; r < R; ++r) pcb[r] = ldp<uint8_t>(a.m.id, pz >> kLogES, off[r] >> kLogES);
    } else if constexpr (!kFull) {
#pragma unroll
      for (int r = 0; r < R; ++r) pcb[r] = ldp<float>(a.m.cb, pz * kCbMul, off[r] * kCbMul);
    }
  };
  // Single-problem tiles load the material ids (or Cb) first: the compiler
  // moves their registers at the loop's back edge, which needs those loads
  // complete, and the vector-memory counter completes in issue order, so the
  // field loads issued after them (and this position's stores) stay in
  // flight there. Issued last, they made that wait a wait for every
  // outstanding access (binary16: vmcnt(0) at the back edge): 512^3
  // binary16 286.7 -> 298.6, 1024^3 314.2 -> 321.9, FP32 236.9 -> 241.0
  // Gcell/s. The batch kernel's tiles keep the material last (256-instance
  // batch 214.4 against 210.9 with it first).
  constexpr bool kMaterialFirst = kSingle;
  auto prefetch = [&](int pp) {
    const size_t pz = (size_t)min(pp, nz - 1) * plane_bytes;
    if constexpr (kMaterialFirst) prefetch_material(pz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        pe[r][c] = ldp<S>(a.src.c[c], pz, off[r]);
        ph[r][c] = ldp<S>(a.src.c[3 + c], pz, off[r]);
      }
    }
    phmz = ldp<S>(a.src.c[5], pz, offm);
    phmx = ldp<S>(a.src.c[3], pz, offm);
    if constexpr (!kMaterialFirst) prefetch_material(pz);
  };
  prefetch(p_begin);
  // With 4-byte fields the loop ends each position with 6R stores. Issuing
  // 6R stores that the range check drops here as well makes the loop's entry
  // look like its back edge: on both paths the prefetched plane is followed
  // by 6R stores, so the loop's wait for that plane never includes the
  // stores just issued.
  if constexpr (kDropStores) {
#pragma unroll
    for (int q = 0; q < 6 * R; ++q)
      stp_drop<S>(a.dst.c[q % 6], 0, (unsigned)plane_bytes, kStoreDrop, S{});
  }
  __syncthreads();

  for (int p = p_begin; p <= p_end; ++p) {
    // ---- level 0: plane p (prefetched), then start loading plane p+1
    float hc[R][3], ec[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ec[r][c] = C::dec(pe[r][c], sc.s[c]);
        hc[r][c] = C::dec(ph[r][c], sc.s[3 + c]);
      }
    float hmz = C::dec(phmz, sc.s[5]), hmx = C::dec(phmx, sc.s[3]);  // row gy0-1
    if constexpr (kIdx) {
#pragma unroll
      for (int r = 0; r < R; ++r) idq[r] = (idq[r] << 8) | (unsigned)pcb[r];
    } else if constexpr (!kFull) {
#pragma unroll
      for (int b = B - 1; b > 0; --b)
#pragma unroll
        for (int r = 0; r < R; ++r) cbq[b][r] = cbq[b - 1][r];
#pragma unroll
      for (int r = 0; r < R; ++r) cbq[0][r] = pcb[r];
    }
    prefetch(p + 1);

#pragma unroll
    for (int b = 1; b <= B; ++b) {
      const int zE = p - b + 1;
      const int zH = p - b;
      const int n = a.n0 + b - 1;
      const bool intzE = zE >= 1 && zE <= nz - 2;
      const bool intzH = zH >= 1 && zH <= nz - 2;

      // ---- E of step n on plane zE from H(level b-1) on zE (hc) and zE-1 (hq[b-1])
      float en[R][3];
      {
        const size_t oE = (size_t)clampi(zE, 0, nz - 1) * plane_bytes;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float hz_xm = from_lane_below(hc[r][2]);
          const float hy_xm = from_lane_below(hc[r][1]);
          const float hz_ym = r > 0 ? hc[r - 1][2] : hmz;
          const float hx_ym = r > 0 ? hc[r - 1][0] : hmx;
          float cr[3];
          cr[0] = (hc[r][2] - hz_ym) - (hc[r][1] - hq[b - 1][r][1]);
          cr[1] = (hc[r][0] - hq[b - 1][r][0]) - (hc[r][2] - hz_xm);
          cr[2] = (hc[r][1] - hy_xm) - (hc[r][0] - hx_ym);
          if (wave_src) apply_sources_row(a.s, gx, gy0 + r, zE, n, dx, 0, cr);
          const bool upd = upd_xy[r] && intzE;
          const float cx = cr[0], cy = cr[1], cz = cr[2];
          if (kFull) {
            const unsigned o = off[r];
            const float ca0 = ldp<float>(a.m.c[0], oE, o);
            const float cb0 = ldp<float>(a.m.c[1], oE, o);
            const float ca1 = ldp<float>(a.m.c[2], oE, o);
            const float cb1 = ldp<float>(a.m.c[3], oE, o);
            const float ca2 = ldp<float>(a.m.c[4], oE, o);
            const float cb2 = ldp<float>(a.m.c[5], oE, o);
            en[r][0] = upd ? ca0 * ec[r][0] + cb0 * cx : ec[r][0];
            en[r][1] = upd ? ca1 * ec[r][1] + cb1 * cy : ec[r][1];
            en[r][2] = upd ? ca2 * ec[r][2] + cb2 * cz : ec[r][2];
          } else {
            float cb;
            if constexpr (kIdx)
              cb = lds_cb[(idq[r] >> (8 * (b - 1))) & 0xffu];
            else
              cb = cbq[b - 1][r];
            en[r][0] = upd ? ec[r][0] + cb * cx : ec[r][0];
            en[r][1] = upd ? ec[r][1] + cb * cy : ec[r][1];
            en[r][2] = upd ? ec[r][2] + cb * cz : ec[r][2];
          }
        }
      }

      // ---- H of step n on plane zH from E(level b) on zH (eq[b]) and zH+1 (en)
      float hn[R][3];
      {
        // read unconditionally (the last wave reads its own row) and select,
        // so that the level stays one basic block and the lane shifts fold
        // into their subtractions
        const int wn = w + 1 < W ? w + 1 : w;
        const float ez_n = lds_e[b - 1][zH & 1][wn][0][lane];
        const float ex_n = lds_e[b - 1][zH & 1][wn][1][lane];
        const float ez_last = w + 1 < W ? ez_n : 0.0f;
        const float ex_last = w + 1 < W ? ex_n : 0.0f;
        const size_t oH = (size_t)clampi(zH, 0, nz - 1) * plane_bytes;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const float ez_xp = from_lane_above(eq[b][r][2]);
          const float ey_xp = from_lane_above(eq[b][r][1]);
          const float ez_yp = r + 1 < R ? eq[b][r + 1][2] : ez_last;
          const float ex_yp = r + 1 < R ? eq[b][r + 1][0] : ex_last;
          float cr[3];
          cr[0] = (en[r][1] - eq[b][r][1]) - (ez_yp - eq[b][r][2]);
          cr[1] = (ez_xp - eq[b][r][2]) - (en[r][0] - eq[b][r][0]);
          cr[2] = (ex_yp - eq[b][r][0]) - (ey_xp - eq[b][r][1]);
          if (wave_src) apply_sources_row(a.s, gx, gy0 + r, zH, n, dx, 3, cr);
          const bool upd = upd_xy[r] && intzH;
          const float cx = cr[0], cy = cr[1], cz = cr[2];
          if (kFull) {
            const unsigned o = off[r];
            const float da0 = ldp<float>(a.m.c[6], oH, o);
            const float db0 = ldp<float>(a.m.c[7], oH, o);
            const float da1 = ldp<float>(a.m.c[8], oH, o);
            const float db1 = ldp<float>(a.m.c[9], oH, o);
            const float da2 = ldp<float>(a.m.c[10], oH, o);
            const float db2 = ldp<float>(a.m.c[11], oH, o);
            hn[r][0] = upd ? da0 * hq[b - 1][r][0] + db0 * cx : hq[b - 1][r][0];
            hn[r][1] = upd ? da1 * hq[b - 1][r][1] + db1 * cy : hq[b - 1][r][1];
            hn[r][2] = upd ? da2 * hq[b - 1][r][2] + db2 * cz : hq[b - 1][r][2];
          } else {
            const float db = a.m.db;
            hn[r][0] = upd ? hq[b - 1][r][0] + db * cx : hq[b - 1][r][0];
            hn[r][1] = upd ? hq[b - 1][r][1] + db * cy : hq[b - 1][r][1];
            hn[r][2] = upd ? hq[b - 1][r][2] + db * cz : hq[b - 1][r][2];
          }
        }
      }

      // ---- DFT: owners store the values after step n on the monitored plane;
      // the FP64 accumulation runs after the launch (dft_snapshot_kernel).
      if (wave_dft) {
        const bool okE = zE >= z_lo && zE < z_hi && (a.d.axis != 2 || zE == a.d.index);
        const bool okH = zH >= z_lo && zH < z_hi && (a.d.axis != 2 || zH == a.d.index);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (okE && ((a.d.mask >> c) & 1u)) {
            const int slot = __builtin_popcount(a.d.mask & ((1u << c) - 1u));
            float* sp = a.d.snap +
                        ((size_t)(b - 1 + a.d.snap_base) * a.d.ncomp + slot) * (size_t)a.d.plane;
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (dft_row(r)) gptr(sp)[dft_base(r) + zE * dft_zmul] = en[r][c];
          }
          if (okH && ((a.d.mask >> (3 + c)) & 1u)) {
            const int slot = __builtin_popcount(a.d.mask & ((1u << (3 + c)) - 1u));
            float* sp = a.d.snap +
                        ((size_t)(b - 1 + a.d.snap_base) * a.d.ncomp + slot) * (size_t)a.d.plane;
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (dft_row(r)) gptr(sp)[dft_base(r) + zH * dft_zmul] = hn[r][c];
          }
        }
      }

      // ---- publish boundary rows for the neighbouring waves
      lds_e[b - 1][zE & 1][w][0][lane] = en[0][2];
      lds_e[b - 1][zE & 1][w][1][lane] = en[0][0];
      if (b < B) {
        lds_h[b - 1][p & 1][w][0][lane] = hn[R - 1][2];
        lds_h[b - 1][p & 1][w][1][lane] = hn[R - 1][0];
      }

      // ---- final level: owned cells go to the destination copy. With
      // 4-byte fields (single-problem launches) every position issues the
      // same 6R stores, and lanes that own no cell of the plane are dropped
      // by the buffer range check (kStoreDrop); 16-bit fields store only
      // owned rows (measured faster for them: 281 against 253 Gcell/s at
      // 512^3, B = 3), and so do the batch's tiles.
      if constexpr (kDropStores) {
        if (b == B) {
          const unsigned dropE = zE >= z_lo && zE < z_hi ? 0u : kStoreDrop;
          const unsigned dropH = zH >= z_lo && zH < z_hi ? 0u : kStoreDrop;
          const size_t oE = (size_t)clampi(zE, 0, nz - 1) * plane_bytes;
          const size_t oH = (size_t)clampi(zH, 0, nz - 1) * plane_bytes;
#pragma unroll
          for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
              stp_drop<S>(a.dst.c[c], oE, (unsigned)plane_bytes, soff[r] | dropE,
                          C::enc(en[r][c], sc.inv[c]));
#pragma unroll
            for (int c = 0; c < 3; ++c)
              stp_drop<S>(a.dst.c[3 + c], oH, (unsigned)plane_bytes, soff[r] | dropH,
                          C::enc(hn[r][c], sc.inv[3 + c]));
          }
        }
      } else if (b == B && own_any) {
        if (zE >= z_lo && zE < z_hi) {
          const size_t oE = (size_t)zE * plane_bytes;
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (own[r]) {
#pragma unroll
              for (int c = 0; c < 3; ++c)
                stp<S>(a.dst.c[c], oE, off[r], C::enc(en[r][c], sc.inv[c]));
            }
        }
        if (zH >= z_lo && zH < z_hi) {
          const size_t oH = (size_t)zH * plane_bytes;
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (own[r]) {
#pragma unroll
              for (int c = 0; c < 3; ++c)
                stp<S>(a.dst.c[3 + c], oH, off[r], C::enc(hn[r][c], sc.inv[3 + c]));
            }
        }
      }

      // ---- advance: level b+1 consumes H(b) on zE-1 (hn), H(b) on zE-2
      // (hq[b], still the previous plane) and E(b) on zE-1 (eq[b] before the
      // shift).
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float e_prev = eq[b][r][c];
          eq[b][r][c] = en[r][c];
          hq[b - 1][r][c] = hc[r][c];
          hc[r][c] = hn[r][c];
          ec[r][c] = e_prev;
        }
      if (b < B || B == 1) __syncthreads();
      if (b < B) {
        const int wp = w > 0 ? w - 1 : 0;
        const float hz_p = lds_h[b - 1][p & 1][wp][0][lane];
        const float hx_p = lds_h[b - 1][p & 1][wp][1][lane];
        hmz = w > 0 ? hz_p : 0.0f;
        hmx = w > 0 ? hx_p : 0.0f;
      }
    }
  }
}

// Copies a wave-uniform object from global memory into registers through
// readfirstlane (a scalar value for every word the caller uses).
template <class T>
__device__ __forceinline__ void uniform_copy(T* dst, const T* src) {
  static_assert(sizeof(T) % 4 == 0, "word-sized objects only");
  const int* s = reinterpret_cast<const int*>(src);
  int* d = reinterpret_cast<int*>(dst);
#pragma unroll
  for (unsigned q = 0; q < sizeof(T) / 4; ++q) d[q] = __builtin_amdgcn_readfirstlane(*gptr(s + q));
}

// Tile of this workgroup: launch position xcd_tile(blockIdx) in tile order,
// or, with a list, the tile whose id (tx + ntx*(ty + nty*tz)) the list holds
// at that position. Lists are in tile order, so each XCD still takes a
// contiguous run of neighbouring tiles.
__device__ __forceinline__ void launch_tile(const ZArgs& a, const int* list, int& tx, int& ty,
                                            int& tz) {
  const int pos = xcd_tile(blockIdx.x, gridDim.x);
  if (list) {
    const int id = list[pos];
    tx = id % a.ntx;
    ty = (id / a.ntx) % a.nty;
    tz = id / (a.ntx * a.nty);
  } else {
    tile_coords(a, pos, tx, ty, tz);
  }
}

// Every tile of the launch (list null) or the listed tiles, with the source
// and DFT code.
template <int B, int R, int W, int kMat, class S>
__global__ __launch_bounds__(64 * W, zstream_waves_per_simd<R>()) void zstream_kernel(
    ZArgs a, const int* __restrict__ list) {
  int tx, ty, tz;
  launch_tile(a, list, tx, ty, tz);
  zstream_tile<B, R, W, kMat, S, true>(a, tx, ty, tz);
}

// The same without the source and DFT code, for tiles that need neither.
// A negative list entry -1-q names packed workgroup q (sub-tiles, see
// zstream_tile).
template <int B, int R, int W, int kMat, class S>
__global__ __launch_bounds__(64 * W, zstream_waves_per_simd<R>()) void zstream_lean_kernel(
    ZArgs a, const int* __restrict__ list, const ZPack* __restrict__ packs) {
  const int pos = xcd_tile(blockIdx.x, gridDim.x);
  const int item = list ? list[pos] : 0;
  int tx, ty, tz, G = 64, WH = W;
  bool active = true;
  if (list && item < 0) {
    ZPack pk;
    uniform_copy(&pk, packs + (-1 - item));
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    G = pk.g;
    WH = W / pk.h;
    const int gi = lane / G, hi = w / WH;
    tx = pk.tx[0][0];
    ty = pk.ty[0][0];
    active = false;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (gi == q && hi == h) {
          tx = pk.tx[q][h];
          ty = pk.ty[q][h];
          active = (pk.own >> (2 * q + h)) & 1;
        }
    tz = pk.tz;
  } else {
    launch_tile(a, list, tx, ty, tz);
  }
  zstream_tile<B, R, W, kMat, S, false>(a, tx, ty, tz, G, WH, active);
}


// Persistent batch kernel: workgroups take (instance, tile) items from the
// launch's queue until it runs dry (every wave of a workgroup sees the same
// item, so all of them leave the loop together). Items of the largest
// instances come first, so the tail of the launch is made of small tiles.
// Items whose tile holds a source cell run the tile with the point-source code,
// all others the build without it: both loops are in the kernel, the item's
// flag (uniform) picks one, and each keeps its own register state.
template <int B, int R, int W, int kMat, class S>
__global__ __launch_bounds__(64 * W, zstream_waves_per_simd<R>()) void zstream_batch_kernel(
    BatchLaunch L) {
  __shared__ int s_item;
  const int n0 = *L.step + L.step_offset;
  const int par = (*L.parity + L.launch_index) & 1;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(L.counter, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    // every wave has read the item before the next one is taken
    __syncthreads();
    if (it >= L.nitems) break;
    // the item and the instance descriptor are uniform: copy them word by
    // word through readfirstlane so that they live in SGPRs
    BatchItem wi;
    uniform_copy(&wi, L.items + it);
    ZArgs a;
    uniform_copy(&a, L.inst + 2 * wi.inst + par);
    a.n0 = n0;
    int tx, ty, tz;
    tile_coords(a, wi.tile, tx, ty, tz);
    // batch form of the tile (kSingle = false): with the batch's small tiles
    // and short streams the fixed-count stores measured slower (256-instance
    // batch 199-207 against 213 Gcell/s)
    if (wi.full)
      zstream_tile<B, R, W, kMat, S, true, false>(a, tx, ty, tz);
    else
      zstream_tile<B, R, W, kMat, S, false, false>(a, tx, ty, tz);
  }
}

__global__ __launch_bounds__(256) void batch_advance_kernel(int* step, int nsteps, int* parity,
                                                            const int* flip, int ngroups,
                                                            unsigned* counters, int ncounters) {
  const int t = threadIdx.x;
  if (t == 0) *step += nsteps;
  for (int g = t; g < ngroups; g += blockDim.x) parity[g] ^= flip[g];
  for (int c = t; c < ncounters; c += blockDim.x) counters[c] = 0u;
}

template <int B, int R, int W, int kMat>
void launch_batch_k(const BatchLaunch& L, int nblocks, hipStream_t stream) {
  hipLaunchKernelGGL((zstream_batch_kernel<B, R, W, kMat, float>), dim3(nblocks), dim3(64 * W), 0,
                     stream, L);
}

template <int B, int R, int W>
const void* batch_kernel_fn(MatKind mk) {
  return mk == kMatIdx
             ? reinterpret_cast<const void*>(&zstream_batch_kernel<B, R, W, kMatIdx, float>)
             : reinterpret_cast<const void*>(&zstream_batch_kernel<B, R, W, kMatIso, float>);
}

template <int R, int W>
const void* batch_kernel_rw(int blt, MatKind mk) {
  switch (blt) {
    case 1: return batch_kernel_fn<1, R, W>(mk);
    case 2: return batch_kernel_fn<2, R, W>(mk);
    case 3: return batch_kernel_fn<3, R, W>(mk);
    case 4: return batch_kernel_fn<4, R, W>(mk);
    default: return nullptr;
  }
}

const void* batch_kernel(const ZTiling& t, MatKind mk) {
  if (t.rows == 2 && t.waves == 4) return batch_kernel_rw<2, 4>(t.blt, mk);
  if (t.rows == 2 && t.waves == 8) return batch_kernel_rw<2, 8>(t.blt, mk);
  if (t.rows == 4 && t.waves == 4) return batch_kernel_rw<4, 4>(t.blt, mk);
  if (t.rows == 4 && t.waves == 8) return batch_kernel_rw<4, 8>(t.blt, mk);
  return nullptr;
}

template <int R, int W>
hipError_t launch_batch_rw(int blt, MatKind mk, const BatchLaunch& L, int nblocks,
                           hipStream_t stream) {
  const bool idx = mk == kMatIdx;
  switch (blt) {
    case 1: idx ? launch_batch_k<1, R, W, kMatIdx>(L, nblocks, stream)
                : launch_batch_k<1, R, W, kMatIso>(L, nblocks, stream); break;
    case 2: idx ? launch_batch_k<2, R, W, kMatIdx>(L, nblocks, stream)
                : launch_batch_k<2, R, W, kMatIso>(L, nblocks, stream); break;
    case 3: idx ? launch_batch_k<3, R, W, kMatIdx>(L, nblocks, stream)
                : launch_batch_k<3, R, W, kMatIso>(L, nblocks, stream); break;
    case 4: idx ? launch_batch_k<4, R, W, kMatIdx>(L, nblocks, stream)
                : launch_batch_k<4, R, W, kMatIso>(L, nblocks, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch(const ZTiling& t, MatKind mk, const BatchLaunch& L, int nblocks,
                        hipStream_t stream) {
  if (t.rows == 2 && t.waves == 4) return launch_batch_rw<2, 4>(t.blt, mk, L, nblocks, stream);
  if (t.rows == 2 && t.waves == 8) return launch_batch_rw<2, 8>(t.blt, mk, L, nblocks, stream);
  if (t.rows == 4 && t.waves == 4) return launch_batch_rw<4, 4>(t.blt, mk, L, nblocks, stream);
  return launch_batch_rw<4, 8>(t.blt, mk, L, nblocks, stream);
}

template <int B, int R, int W, int kMat, class S>
void launch_k(const ZArgs& a, const ZPart& part, hipStream_t stream) {
  const unsigned n = part.list ? (unsigned)part.nlist
                               : (unsigned)((long long)a.ntx * a.nty * a.ntz);
  if (part.form == kZLean) {
    // the lean form is built for the planner's 4 x 8 tile only
    if constexpr (R == 4 && W == 8)
      hipLaunchKernelGGL((zstream_lean_kernel<B, R, W, kMat, S>), dim3(n), dim3(64 * W), 0,
                         stream, a, part.list, part.packs);
  } else {
    hipLaunchKernelGGL((zstream_kernel<B, R, W, kMat, S>), dim3(n), dim3(64 * W), 0, stream, a,
                       part.list);
  }
}

template <int B, int R, int W>
hipError_t launch_one(const ZArgs& a, Storage st, MatKind mk, const ZPart& part,
                      hipStream_t stream) {
  if (st == kF16) {
    if (mk == kMatIdx)
      launch_k<B, R, W, kMatIdx, _Float16>(a, part, stream);
    else
      launch_k<B, R, W, kMatIso, _Float16>(a, part, stream);
  } else if (st == kBF16) {
    if (mk == kMatIdx)
      launch_k<B, R, W, kMatIdx, __bf16>(a, part, stream);
    else
      launch_k<B, R, W, kMatIso, __bf16>(a, part, stream);
  } else if (mk == kMatFull) {
    launch_k<B, R, W, kMatFull, float>(a, part, stream);
  } else if (mk == kMatIdx) {
    launch_k<B, R, W, kMatIdx, float>(a, part, stream);
  } else {
    launch_k<B, R, W, kMatIso, float>(a, part, stream);
  }
  return hipGetLastError();
}

template <int R, int W>
hipError_t launch_rw(int blt, const ZArgs& a, Storage st, MatKind mk, const ZPart& part,
                     hipStream_t stream) {
  switch (blt) {
    case 1: return launch_one<1, R, W>(a, st, mk, part, stream);
    case 2: return launch_one<2, R, W>(a, st, mk, part, stream);
    case 3: return launch_one<3, R, W>(a, st, mk, part, stream);
    case 4: return launch_one<4, R, W>(a, st, mk, part, stream);
    default: return hipErrorInvalidValue;
  }
}

__global__ __launch_bounds__(256) void dft_snapshot_kernel(Dft d, int nlev, int run_step0) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= d.plane) return;
  for (int c = 0; c < 6; ++c) {
    if (!((d.mask >> c) & 1u)) continue;
    const int slot = __builtin_popcount(d.mask & ((1u << c) - 1u));
    for (int f = 0; f < d.nf; ++f) {
      const long long o = ((long long)slot * d.nf + f) * d.plane + p;
      double re = d.re[o], im = d.im[o];
      for (int b = 0; b < nlev; ++b) {
        const double v = (double)d.snap[((long long)b * d.ncomp + slot) * d.plane + p];
        re = re + v * d.wc[(long long)f * d.run_len + run_step0 + b];
        im = im - v * d.ws[(long long)f * d.run_len + run_step0 + b];
      }
      d.re[o] = re;
      d.im[o] = im;
    }
  }
}

}  // namespace

hipError_t launch_dft_snapshot(const Dft& d, int nlev, int run_step0, hipStream_t stream) {
  const unsigned blocks = (unsigned)((d.plane + 255) / 256);
  hipLaunchKernelGGL(dft_snapshot_kernel, dim3(blocks), dim3(256), 0, stream, d, nlev, run_step0);
  return hipGetLastError();
}

bool ztiling_supported(const ZTiling& t, Storage st, MatKind mk) {
  if (st == kF64 || ((st == kF16 || st == kBF16) && mk == kMatFull)) return false;
  if (t.blt < 1 || t.blt > 4 || t.zchunk < 1) return false;
  const bool rw = (t.rows == 2 && t.waves == 4) || (t.rows == 2 && t.waves == 8) ||
                  (t.rows == 4 && t.waves == 4) || (t.rows == 4 && t.waves == 8);
  if (!rw) return false;
  return t.rows * t.waves - 2 * t.blt >= 1 && 64 - 2 * t.blt >= 1;
}

hipError_t launch_zstream_batch(const ZTiling& t, MatKind mk, const BatchLaunch& L, int nblocks,
                                hipStream_t stream) {
  if (!ztiling_supported(t, kF32, mk) || mk == kMatFull || nblocks < 1)
    return hipErrorInvalidValue;
  return launch_batch(t, mk, L, nblocks, stream);
}

int zstream_batch_blocks(const ZTiling& t, MatKind mk, int device) {
  const void* fn = batch_kernel(t, mk);
  if (!fn) return 0;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * t.waves, 0) != hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return 0;
  return per_cu * cus;
}

hipError_t launch_batch_advance(int* step, int nsteps, int* parity, const int* flip, int ngroups,
                                unsigned* counters, int ncounters, hipStream_t stream) {
  hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(256), 0, stream, step, nsteps, parity,
                     flip, ngroups, counters, ncounters);
  return hipGetLastError();
}

bool zstream_lean_supported(const ZTiling& t) { return t.rows == 4 && t.waves == 8; }

// Column width of the tile order (ZArgs::scw) for a storage format.
int zstream_scw(Storage st) { return st == kF32 ? 10 : 0; }

namespace {
// Does tile (tx, ty, tz) need the source or DFT code: a source cell in its
// input window (x [x0, x0+64), y [y0, y0+rows*waves)) and in the planes its
// stream computes (z_lo-2B .. z_hi-1+B), or owned cells on the monitored
// plane?
bool tile_needs_full(const ZTiling& t, const Grid& g, const SourceDesc* src, int nsrc,
                     const Dft& d, int tx, int ty, int tz) {
  const int B = t.blt;
  const int ox = 64 - 2 * B, oy = t.rows * t.waves - 2 * B;
  const int x0 = tx * ox - B, y0 = ty * oy - B;
  int z_lo, z_hi;
  chunk_range(t.zchunk, g.nz, t.zsplit, tz, z_lo, z_hi);
  for (int q = 0; q < nsrc; ++q)
    if (src[q].i >= x0 && src[q].i < x0 + 64 && src[q].j >= y0 &&
        src[q].j < y0 + t.rows * t.waves && src[q].k >= z_lo - 2 * B &&
        src[q].k <= z_hi - 1 + B)
      return true;
  if (d.ncomp > 0) {
    if (d.axis == 2) return d.index >= z_lo && d.index < z_hi;
    if (d.axis == 0) return d.index >= tx * ox && d.index < tx * ox + ox;
    return d.index >= ty * oy && d.index < ty * oy + oy;
  }
  return false;
}
}  // namespace

void zstream_classify(const ZTiling& t, int scw, const Grid& g, const SourceDesc* src, int nsrc,
                      const Dft& d, std::vector<int>& extra, std::vector<int>& lean) {
  ZArgs a{};
  ztile_counts(t, g.nx, g.ny, g.nz, a.ntx, a.nty, a.ntz);
  a.scw = scw;
  extra.clear();
  lean.clear();
  const int ntiles = a.ntx * a.nty * a.ntz;
  for (int pos = 0; pos < ntiles; ++pos) {
    // the kernels' tile order, so that both lists keep neighbouring tiles
    // together
    int tx, ty, tz;
    tile_coords(a, pos, tx, ty, tz);
    (tile_needs_full(t, g, src, nsrc, d, tx, ty, tz) ? extra : lean)
        .push_back(tx + a.ntx * (ty + a.nty * tz));
  }
}

// Workgroups one XCD runs at once with the 4 x 8 tile: 32 CUs, one workgroup
// of 8 waves per CU (2 waves per SIMD).
constexpr int kXcdSlots = 32;

void zstream_chunk_groups(const ZTiling& t, int ntz, int ncol, std::vector<int>& lean,
                          const std::vector<ZPack>& packs) {
  // items of each z chunk, in the base order
  std::vector<std::vector<int>> chunk(ntz);
  for (const int it : lean) chunk[it < 0 ? packs[-1 - it].tz : it / ncol].push_back(it);
  size_t longest = 0;
  for (const auto& v : chunk) longest = std::max(longest, v.size());
  const int D = t.zchunk, B = t.blt;
  const size_t gc = (size_t)std::max(1, (kXcdSlots * D + (D + 2 * B) / 2) / (D + 2 * B));
  lean.clear();
  for (size_t g0 = 0; g0 < longest; g0 += gc)
    for (int tz = 0; tz < ntz; ++tz)
      for (size_t i = g0; i < std::min(chunk[tz].size(), g0 + gc); ++i) lean.push_back(chunk[tz][i]);
}

void zstream_plan_items(const ZTiling& t, Storage st, const Grid& g, const SourceDesc* src,
                        int nsrc, const Dft& d, std::vector<int>& extra, std::vector<int>& lean,
                        std::vector<ZPack>& packs) {
  const int scw = zstream_scw(st);
  // chunk groups for FP32 fields only: binary16 launches (64-plane chunks,
  // a halo of 6 planes in 70) measured 1.6-2.4 % slower with them (768^3,
  // 1024^3) and keep the plain order
  const bool grouped = st == kF32;
  packs.clear();
  ZArgs a{};
  ztile_counts(t, g.nx, g.ny, g.nz, a.ntx, a.nty, a.ntz);
  const int B = t.blt, W = t.waves, R = t.rows;
  const int ox = 64 - 2 * B, oy = R * W - 2 * B;
  // remainder column and row: cells of the last tile column / row
  const int rx = g.nx - (a.ntx - 1) * ox, ry = g.ny - (a.nty - 1) * oy;
  int G = 64, H = 1;
  if (scw <= 0 || a.ntx <= scw) {  // plain x-fastest order
    if (a.ntx > 1 && rx + 2 * B <= 16)
      G = 16;
    else if (a.ntx > 1 && rx + 2 * B <= 32)
      G = 32;
    if (a.nty > 1 && W % 2 == 0 && ry + 2 * B <= (W / 2) * R) H = 2;
  }
  if (G == 64 && H == 1) {
    zstream_classify(t, scw, g, src, nsrc, d, extra, lean);
    if (grouped) zstream_chunk_groups(t, a.ntz, a.ntx * a.nty, lean, packs);
    return;
  }
  extra.clear();
  lean.clear();
  // pending sub-tiles of the remainder column (x packs: 64/G per workgroup,
  // full height) and row (y packs: 2 per workgroup, half height)
  std::vector<std::pair<int, int>> xp, yp;
  auto emit = [&](std::vector<std::pair<int, int>>& pend, int tz, bool xpack) {
    if (pend.empty()) return;
    ZPack pk{};
    pk.tz = tz;
    pk.g = xpack ? G : 64;
    pk.h = xpack ? 1 : 2;
    pk.own = 0;
    const int slots = xpack ? 64 / G : 2;
    for (int q = 0; q < 4; ++q)
      for (int k = 0; k < 2; ++k) {
        pk.tx[q][k] = pend[0].first;
        pk.ty[q][k] = pend[0].second;
      }
    for (int i = 0; i < slots && i < (int)pend.size(); ++i) {
      const int q = xpack ? i : 0, k = xpack ? 0 : i;
      pk.tx[q][k] = pend[i].first;
      pk.ty[q][k] = pend[i].second;
      pk.own |= 1 << (2 * q + k);
    }
    lean.push_back(-1 - (int)packs.size());
    packs.push_back(pk);
    pend.clear();
  };
  for (int tz = 0; tz < a.ntz; ++tz) {
    for (int ty = 0; ty < a.nty; ++ty)
      for (int tx = 0; tx < a.ntx; ++tx) {
        const int id = tx + a.ntx * (ty + a.nty * tz);
        if (tile_needs_full(t, g, src, nsrc, d, tx, ty, tz)) {
          extra.push_back(id);
        } else if (G < 64 && tx == a.ntx - 1) {
          // narrow tile: its G lanes cover the rx remainder cells plus halos
          xp.emplace_back(tx, ty);
          if ((int)xp.size() == 64 / G) emit(xp, tz, true);
        } else if (H == 2 && ty == a.nty - 1) {
          // short tile: W/2 waves cover the ry remainder rows plus halos
          yp.emplace_back(tx, ty);
          if ((int)yp.size() == 2) emit(yp, tz, false);
        } else {
          lean.push_back(id);
        }
      }
    emit(xp, tz, true);
    emit(yp, tz, false);
  }
  if (grouped) zstream_chunk_groups(t, a.ntz, a.ntx * a.nty, lean, packs);
}

hipError_t launch_zstream(const ZTiling& t, Storage st, const Fields& src, const Fields& dst,
                          const Scales& sc, const Material& m, MatKind mk, const Grid& g,
                          const Sources& s, const Dft& d, int n0, int run_step0,
                          const ZPart& part, hipStream_t stream) {
  if (!ztiling_supported(t, st, mk)) return hipErrorInvalidValue;
  if (part.form == kZLean && !zstream_lean_supported(t)) return hipErrorInvalidValue;
  if (part.list && part.nlist < 1) return hipErrorInvalidValue;
  // in-plane byte offsets are 32-bit, and a store offset with bit 31 set is
  // dropped (stp_drop): every plane stays below 2^31 bytes
  if ((unsigned long long)g.sxy * 4ull >= (1ull << 31)) return hipErrorInvalidValue;
  ZArgs a;
  a.src = src;
  a.dst = dst;
  a.sc = sc;
  a.m = m;
  a.g = g;
  a.s = s;
  a.d = d;
  a.n0 = n0;
  a.run_step0 = run_step0;
  ztile_counts(t, g.nx, g.ny, g.nz, a.ntx, a.nty, a.ntz);
  a.zc = t.zchunk;
  a.zsplit = t.zsplit;
  // FP32 on rows of more than 10 tiles: columns 10 tiles wide keep the tiles
  // an XCD runs together in a compact block that shares halo rows in L2
  // (1024^3: ~200 instead of ~190 Gcell/s); 16-bit launches measured no gain
  a.scw = zstream_scw(st);
  if (t.rows == 2 && t.waves == 4) return launch_rw<2, 4>(t.blt, a, st, mk, part, stream);
  if (t.rows == 2 && t.waves == 8) return launch_rw<2, 8>(t.blt, a, st, mk, part, stream);
  if (t.rows == 4 && t.waves == 4) return launch_rw<4, 4>(t.blt, a, st, mk, part, stream);
  return launch_rw<4, 8>(t.blt, a, st, mk, part, stream);
}

}  // namespace gd
