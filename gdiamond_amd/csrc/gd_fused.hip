// FUSED stepping kernel for gfx950: one launch per step, E then H in one pass.
//
// Same update and results as the PLAIN pair (E launch, then H launch: the
// reference's update_FDTD_gpu_3D_warp_underutilization_fix,
// gdiamond_gpu.cuh:896-913, with updateE_3Dmap_fix / updateH_3Dmap_fix,
// kernels.cuh:337-394), fused into one pass over the grid that reads the
// fields of step n from one copy and writes step n+1 to the other: 52 bytes
// per cell-update for FP32 fields with a dense Cb (49 with the material id),
// against 76 for the in-place pair, which reads E and H twice.
//
// Geometry: a wave owns R consecutive y rows of a 62-cell x segment and a
// chunk of z planes; lanes 1..62 own x cells, lanes 0 and 63 are the x-1 and
// x+1 halo, reached by DPP lane shifts. H of step n+1 at (i,j,k) needs E of
// step n+1 at i+1, j+1 and k+1, so a wave also computes E on the row above
// its rows (Ex and Ez, the components H's y-derivative reads) and on the
// plane after its chunk, from the same inputs with the same operations as
// the wave that owns those cells: no wave waits for another, and the kernel
// needs no LDS exchange and no barrier. Each wave marches along z; H runs
// one plane behind E, and each plane's inputs are loaded one plane ahead.
// Many short, light waves per CU keep enough loads in flight to cover the
// memory latency, as in the PLAIN kernels.
//
// Physical boundary cells are copied unchanged, exactly as the reference
// updates only 1..n-2 (gdiamond.hpp:1194-1221). The DFT of E and of H is
// accumulated in place, as in the PLAIN kernels, from the values after the
// step as stored (16-bit storage rounds every step, as PLAIN does).
#include <type_traits>

#include "gd_common.hpp"

namespace gd {

namespace {

// Three rows per wave, one wave per workgroup, 8-plane z chunks: many short,
// light waves whose halo rows and halo planes are their neighbours' own,
// found in cache. At 512^3 (FP32 / binary16 / FP64 Gcell/s, rows x waves x
// planes, measured on two boxes): 3 x 1 x 8 77.3 / 131 / 36.5 (here);
// 3 x 1 x 7 76.3 / 128; 3 x 1 x 9 74.8 / 130; 3 x 2 x 8 75.0 / 126;
// 2 x 1 x 8 75.7 / 125; 4 x 1 x 8 76.8 / 124; 5 x 1 x 8 75.8 / 89;
// 2 x 2 x 8 74.6 / 128; 1 x 4 x 8 70.8 / 123; 1 x 8 x 32 65.4 / 107;
// 4 x 4 x 32 61.2 / 106; 1 x 8 x 64 55.1 / 103 (PLAIN: 58.4 / 63 / 30.3).
constexpr int kFuW = 1;    // waves per workgroup
constexpr int kFuR = 3;    // owned rows per wave
constexpr int kFuKc = 8;   // z planes per wave
constexpr int kFuOx = 62;  // owned x cells per wave (lanes 1..62)

template <class S, int kMat>
__global__ __launch_bounds__(64 * kFuW) void fused_step_kernel(Fields in, Fields out, Scales sc,
                                                               Material m, Grid g, Sources s,
                                                               Dft d, int n, int run_step,
                                                               int nsx, int nsy, int gcol) {
  using C = Codec<S>;
  using T = typename C::T;
  constexpr int R = kFuR;
  constexpr bool kFull = kMat == kMatFull;
  constexpr bool kIdx = kMat == kMatIdx;
  constexpr unsigned ES = sizeof(S);
  // material element: the 1-byte id, or a coefficient in the compute type
  constexpr unsigned MS = kIdx ? 1u : (unsigned)sizeof(T);

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  // Groups of gcol consecutive x-y columns, each walked through every z
  // chunk before the next group (the last group may be narrower).
  const int ncol = nsx * nsy;
  const int nsz = gridDim.x / ncol;
  const int grp = b / (gcol * nsz);
  const int gw = min(gcol, ncol - grp * gcol);
  const int rem = b - grp * gcol * nsz;
  const int col = grp * gcol + rem % gw;
  const int sx = col % nsx;
  const int sy = col / nsx;
  const int sz = rem / gw;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  // Cb of each material id, in LDS (the lookup then costs no dependent
  // global load after the id's own)
  __shared__ float lds_cb[kIdx ? kMaxMaterials : 1];
  if constexpr (kIdx) {
    for (int q = threadIdx.x; q < kMaxMaterials; q += 64 * kFuW) lds_cb[q] = m.table[q];
    __syncthreads();
  }
  const int j0 = (sy * kFuW + w) * R;
  if (j0 >= ny) return;  // the whole wave is beyond the grid

  const int i = sx * kFuOx - 1 + lane;
  const int ci = clampi(i, 0, nx - 1);
  const bool own_x = lane >= 1 && lane <= kFuOx && i < nx;
  const bool int_x = i >= 1 && i <= nx - 2;
  const int k0 = sz * kFuKc;
  const int k1 = min(nz, k0 + kFuKc);
  const size_t plane_cells = (size_t)g.sxy;
  const T dx = (T)g.dx;
  const T db = sizeof(T) == 8 ? (T)m.db64 : (T)m.db;

  // in-plane cell offsets (clamped into the grid) of the own rows, the row
  // below (j0-1) and the row above (j0+R)
  unsigned cell[R];
  bool own[R], upd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    cell[r] = (unsigned)ci + (unsigned)clampi(j, 0, ny - 1) * (unsigned)nx;
    own[r] = own_x && j < ny;
    upd[r] = int_x && j >= 1 && j <= ny - 2;
  }
  const unsigned cell_b = (unsigned)ci + (unsigned)clampi(j0 - 1, 0, ny - 1) * (unsigned)nx;
  const int ja = j0 + R;
  const unsigned cell_a = (unsigned)ci + (unsigned)clampi(ja, 0, ny - 1) * (unsigned)nx;
  const bool upd_a = int_x && ja >= 1 && ja <= ny - 2;

  // E is computed one plane past the chunk when that plane exists
  const int k_end = k1 < nz ? k1 : k1 - 1;

  // Wave-uniform flags for the rare work: does a source sit in one of this
  // wave's columns (any lane, own rows and the row above, planes it
  // computes), does the wave own cells of the monitored plane? The stepping
  // loop runs that code only when they are set. A cell without a source gets
  // no term at all, which equals the reference's x - (+0)*dx == x bit for bit.
  bool any_src = false;
  for (int q = 0; q < s.n; ++q) {
    const SourceDesc sd = s.d[q];
    any_src = any_src || (sd.i == i && sd.j >= j0 && sd.j <= ja && sd.k >= k0 && sd.k <= k_end);
  }
  const bool wave_src = __builtin_amdgcn_ballot_w64(any_src) != 0;
  bool on_plane = false;
  if (d.ncomp > 0) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      on_plane = on_plane || (own[r] && (d.axis == 0 ? i == d.index
                                         : d.axis == 1 ? j0 + r == d.index
                                                       : d.index >= k0 && d.index < k1));
  }
  const bool wave_dft = __builtin_amdgcn_ballot_w64(on_plane) != 0;
  const bool dft_e = wave_dft && (d.mask & 7u) != 0u;
  const bool dft_h = wave_dft && (d.mask & 56u) != 0u;

  auto fld = [&](int c, size_t pz, unsigned q) {
    return C::dec(ldp<S, false>(in.c[c], pz * ES, q * ES), sc.s[c]);
  };
  auto coef = [&](int a, size_t pz, unsigned q) {
    return ldp<T, false>(m.c[a], pz * sizeof(T), q * (unsigned)sizeof(T));
  };

  // H of step n on plane k-1 (own rows; Hy of the row above), and E of step
  // n+1 on plane k-1 (own rows; Ex, Ez of the row above)
  T hp[R][3], ep[R][3], hs[R][3];  // hs: H of step n+1 on plane k-2, to store
  T hy_pa = 0, ex_pa = 0, ez_pa = 0;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      hp[r][c] = 0;
      ep[r][c] = 0;
      hs[r][c] = 0;
    }
  if (k0 >= 1) {
    const size_t pz = (size_t)(k0 - 1) * plane_cells;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      hp[r][0] = fld(3, pz, cell[r]);
      hp[r][1] = fld(4, pz, cell[r]);
    }
    hy_pa = fld(4, pz, cell_a);
  }

  // Inputs of a plane in the storage format, loaded one plane ahead so that
  // their latency overlaps the updates of the current plane: own rows, Hz and
  // Hx of the row below, Ex Ez Hx Hy Hz of the row above, and the material id
  // of the own rows and the row above (isotropic material forms).
  using MT = typename std::conditional<kIdx, uint8_t, T>::type;
  S pe[R][3], ph[R][3], p_hzb, p_hxb, p_exa, p_eza, p_hxa, p_hya, p_hza;
  MT pm[R], pma;
  auto raw = [&](int c, size_t pz, unsigned q) { return ldp<S, false>(in.c[c], pz * ES, q * ES); };
  // The material first: the top of the next iteration looks Cb up in LDS
  // from these ids before it needs every field value, and the vector-memory
  // counter completes in issue order (512^3 binary16 131.3 -> 136.6 Gcell/s,
  // FP32 unchanged).
  auto prefetch = [&](int kk) {
    const size_t pz = (size_t)kk * plane_cells;
    if constexpr (!kFull) {
#pragma unroll
      for (int r = 0; r < R; ++r) pm[r] = ldp<MT, false>(kIdx ? (const void*)m.id : m.cb, pz * MS, cell[r] * MS);
      pma = ldp<MT, false>(kIdx ? (const void*)m.id : m.cb, pz * MS, cell_a * MS);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pe[r][c] = raw(c, pz, cell[r]);
        ph[r][c] = raw(3 + c, pz, cell[r]);
      }
    p_hzb = raw(5, pz, cell_b);
    p_hxb = raw(3, pz, cell_b);
    p_exa = raw(0, pz, cell_a);
    p_eza = raw(2, pz, cell_a);
    p_hxa = raw(3, pz, cell_a);
    p_hya = raw(4, pz, cell_a);
    p_hza = raw(5, pz, cell_a);
  };
  prefetch(k0);

  for (int k = k0; k <= k_end; ++k) {
    const size_t pz = (size_t)k * plane_cells;
    const bool intz = k >= 1 && k <= nz - 2;
    T e[R][3], h[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        e[r][c] = C::dec(pe[r][c], sc.s[c]);
        h[r][c] = C::dec(ph[r][c], sc.s[3 + c]);
      }
    const T hz_b = C::dec(p_hzb, sc.s[5]), hx_b = C::dec(p_hxb, sc.s[3]);
    const T ex_a = C::dec(p_exa, sc.s[0]), ez_a = C::dec(p_eza, sc.s[2]);
    const T hx_a = C::dec(p_hxa, sc.s[3]), hy_a = C::dec(p_hya, sc.s[4]);
    const T hz_a = C::dec(p_hza, sc.s[5]);
    T cbr[R], cba;
    if constexpr (!kFull) {
#pragma unroll
      for (int r = 0; r < R; ++r) cbr[r] = kIdx ? (T)lds_cb[(unsigned)pm[r]] : (T)pm[r];
      cba = kIdx ? (T)lds_cb[(unsigned)pma] : (T)pma;
    }
    // The stores of the previous iteration (E of plane k-1, H of plane k-2)
    // are issued here, before the next plane's loads: the vector-memory
    // counter completes in issue order, so the wait for those loads one
    // iteration from now finds them long done instead of waiting for stores
    // just issued.
    if (k - 1 >= k0) {
      const size_t pe_off = (size_t)(k - 1) * plane_cells * ES;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (own[r]) {
#pragma unroll
          for (int c = 0; c < 3; ++c)
            stp<S, false>(out.c[c], pe_off, cell[r] * ES, C::enc(ep[r][c], sc.inv[c]));
        }
      if (k - 2 >= k0) {
        const size_t ph_off = (size_t)(k - 2) * plane_cells * ES;
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (own[r]) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
              stp<S, false>(out.c[3 + c], ph_off, cell[r] * ES, C::enc(hs[r][c], sc.inv[3 + c]));
          }
      }
    }
    if (k < k_end) prefetch(k + 1);

    // ---- E of step n+1 on plane k, own rows
    T en[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const T hz_jm = r > 0 ? h[r - 1][2] : hz_b;
      const T hx_jm = r > 0 ? h[r - 1][0] : hx_b;
      const T hz_im = from_lane_below(h[r][2]);
      const T hy_im = from_lane_below(h[r][1]);
      T cx = (h[r][2] - hz_jm) - (h[r][1] - hp[r][1]);
      T cy = (h[r][0] - hp[r][0]) - (h[r][2] - hz_im);
      T cz = (h[r][1] - hy_im) - (h[r][0] - hx_jm);
      if (wave_src) {
        float jx, jy, jz;
        source_terms(s, 0, i, j0 + r, k, n, jx, jy, jz);
        cx = cx - (T)jx * dx;
        cy = cy - (T)jy * dx;
        cz = cz - (T)jz * dx;
      }
      const bool u = upd[r] && intz;
      if constexpr (kFull) {
        const T ca0 = coef(0, pz, cell[r]), cb0 = coef(1, pz, cell[r]);
        const T ca1 = coef(2, pz, cell[r]), cb1 = coef(3, pz, cell[r]);
        const T ca2 = coef(4, pz, cell[r]), cb2 = coef(5, pz, cell[r]);
        en[r][0] = u ? ca0 * e[r][0] + cb0 * cx : e[r][0];
        en[r][1] = u ? ca1 * e[r][1] + cb1 * cy : e[r][1];
        en[r][2] = u ? ca2 * e[r][2] + cb2 * cz : e[r][2];
      } else {
        const T cb = cbr[r];
        en[r][0] = u ? e[r][0] + cb * cx : e[r][0];
        en[r][1] = u ? e[r][1] + cb * cy : e[r][1];
        en[r][2] = u ? e[r][2] + cb * cz : e[r][2];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) en[r][c] = C::round(en[r][c], sc.s[c], sc.inv[c]);
    }

    // ---- Ex, Ez of step n+1 on plane k, row above (read by H's j+1 terms)
    T exa_n, eza_n;
    {
      const T hy_im = from_lane_below(hy_a);
      T cx = (hz_a - h[R - 1][2]) - (hy_a - hy_pa);
      T cz = (hy_a - hy_im) - (hx_a - h[R - 1][0]);
      if (wave_src) {
        float jx, jy, jz;
        source_terms(s, 0, i, ja, k, n, jx, jy, jz);
        cx = cx - (T)jx * dx;
        cz = cz - (T)jz * dx;
      }
      const bool u = upd_a && intz;
      if constexpr (kFull) {
        const T ca0 = coef(0, pz, cell_a), cb0 = coef(1, pz, cell_a);
        const T ca2 = coef(4, pz, cell_a), cb2 = coef(5, pz, cell_a);
        exa_n = u ? ca0 * ex_a + cb0 * cx : ex_a;
        eza_n = u ? ca2 * ez_a + cb2 * cz : ez_a;
      } else {
        const T cb = cba;
        exa_n = u ? ex_a + cb * cx : ex_a;
        eza_n = u ? ez_a + cb * cz : ez_a;
      }
      exa_n = C::round(exa_n, sc.s[0], sc.inv[0]);
      eza_n = C::round(eza_n, sc.s[2], sc.inv[2]);
    }

    // ---- DFT of the owned E of plane k (the plane past the chunk belongs to
    // the next); the E values are stored at the top of the next iteration
    if (k < k1) {
      if (dft_e) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const long long p = own[r] ? dft_plane_index(d, g, i, j0 + r, k) : -1;
          if (p >= 0)
            dft_accumulate_group<true>(d, p, (double)en[r][0], (double)en[r][1],
                                       (double)en[r][2], run_step);
        }
      }
    }

    // ---- H of step n+1 on plane k-1 from E of step n+1 on planes k-1, k
    if (k > k0) {
      const int kh = k - 1;
      const size_t ph = (size_t)kh * plane_cells;
      const bool intzh = kh >= 1 && kh <= nz - 2;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T ez_jp = r + 1 < R ? ep[r + 1][2] : ez_pa;
        const T ex_jp = r + 1 < R ? ep[r + 1][0] : ex_pa;
        const T ez_ip = from_lane_above(ep[r][2]);
        const T ey_ip = from_lane_above(ep[r][1]);
        T cx = (en[r][1] - ep[r][1]) - (ez_jp - ep[r][2]);
        T cy = (ez_ip - ep[r][2]) - (en[r][0] - ep[r][0]);
        T cz = (ex_jp - ep[r][0]) - (ey_ip - ep[r][1]);
        if (wave_src) {
          float mx, my, mz;
          source_terms(s, 3, i, j0 + r, kh, n, mx, my, mz);
          cx = cx - (T)mx * dx;
          cy = cy - (T)my * dx;
          cz = cz - (T)mz * dx;
        }
        const bool u = upd[r] && intzh;
        T hn[3];
        if constexpr (kFull) {
          const T da0 = coef(6, ph, cell[r]), db0 = coef(7, ph, cell[r]);
          const T da1 = coef(8, ph, cell[r]), db1 = coef(9, ph, cell[r]);
          const T da2 = coef(10, ph, cell[r]), db2 = coef(11, ph, cell[r]);
          hn[0] = u ? da0 * hp[r][0] + db0 * cx : hp[r][0];
          hn[1] = u ? da1 * hp[r][1] + db1 * cy : hp[r][1];
          hn[2] = u ? da2 * hp[r][2] + db2 * cz : hp[r][2];
        } else {
          hn[0] = u ? hp[r][0] + db * cx : hp[r][0];
          hn[1] = u ? hp[r][1] + db * cy : hp[r][1];
          hn[2] = u ? hp[r][2] + db * cz : hp[r][2];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) hs[r][c] = hn[c];
        if (dft_h) {
          const long long p = own[r] ? dft_plane_index(d, g, i, j0 + r, kh) : -1;
          if (p >= 0)
            dft_accumulate_group<false>(d, p, (double)C::round(hn[0], sc.s[3], sc.inv[3]),
                                        (double)C::round(hn[1], sc.s[4], sc.inv[4]),
                                        (double)C::round(hn[2], sc.s[5], sc.inv[5]), run_step);
        }
      }
    }

#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        hp[r][c] = h[r][c];
        ep[r][c] = en[r][c];
      }
    hy_pa = hy_a;
    ex_pa = exa_n;
    ez_pa = eza_n;
  }

  // Stores of the last iteration: E of plane k_end when it is owned, H of
  // plane k_end-1.
  if (k_end < k1) {
    const size_t pe_off = (size_t)k_end * plane_cells * ES;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (own[r]) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          stp<S, false>(out.c[c], pe_off, cell[r] * ES, C::enc(ep[r][c], sc.inv[c]));
      }
  }
  if (k_end - 1 >= k0) {
    const size_t ph_off = (size_t)(k_end - 1) * plane_cells * ES;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (own[r]) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          stp<S, false>(out.c[3 + c], ph_off, cell[r] * ES, C::enc(hs[r][c], sc.inv[3 + c]));
      }
  }

  // The last plane of the grid is a boundary plane: its H is copied unchanged
  // (hp holds it after the loop).
  if (k1 == nz) {
    const int kh = nz - 1;
    const size_t ph = (size_t)kh * plane_cells;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (own[r]) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          stp<S, false>(out.c[3 + c], ph * ES, cell[r] * ES, C::enc(hp[r][c], sc.inv[3 + c]));
      }
      if (dft_h) {
        const long long p = own[r] ? dft_plane_index(d, g, i, j0 + r, kh) : -1;
        if (p >= 0)
          dft_accumulate_group<false>(d, p, (double)hp[r][0], (double)hp[r][1], (double)hp[r][2],
                                      run_step);
      }
    }
  }
}

template <class S, int kMat>
hipError_t launch_fused(const Fields& in, const Fields& out, const Scales& sc, const Material& m,
                        const Grid& g, const Sources& s, const Dft& d, int n, int run_step,
                        hipStream_t stream) {
  const int nsx = (g.nx + kFuOx - 1) / kFuOx;
  const int nsy = (g.ny + kFuW * kFuR - 1) / (kFuW * kFuR);
  const int nsz = (g.nz + kFuKc - 1) / kFuKc;
  const long long nb = (long long)nsx * nsy * nsz;
  if (nb > INT32_MAX) return hipErrorInvalidValue;
  // Column-group width. Each XCD takes a contiguous run of the launch (the
  // xcd_tile remap) and holds `slots` workgroups at a time, so in a group of
  // gcol columns a z chunk starts gcol/slots of a chunk's stream after the
  // chunk below it; with gcol = slots * D / (D + 2) (D = kFuKc planes, plus
  // one halo plane each side) that is when the chunk below reaches the two
  // planes the two read in common (its last owned plane and its top halo),
  // so both read them at about the same time. 512^3: 205 columns for the
  // 2 waves/SIMD of the FP32 and binary16 builds; against the plain x, y, z
  // order FP32 79.6 -> 82.6 Gcell/s, binary16 130.6 -> 142.8 (at 228; 114 to
  // 456 columns all measured within 1 % of that). A grid of fewer columns
  // than gcol is one group, i.e. the plain order.
  static const int gcol = [] {
    int dev = 0, occ = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fused_step_kernel<S, kMat>,
                                                     64 * kFuW, 0) != hipSuccess ||
        occ < 1 || cus < 8)
      return 1 << 30;  // unknown: one group, the plain order
    const int slots = occ * (cus / 8);
    return max(1, (slots * kFuKc + (kFuKc + 2) / 2) / (kFuKc + 2));
  }();
  hipLaunchKernelGGL((fused_step_kernel<S, kMat>), dim3((unsigned)nb), dim3(64 * kFuW), 0, stream,
                     in, out, sc, m, g, s, d, n, run_step, nsx, nsy, min(gcol, nsx * nsy));
  return hipGetLastError();
}

}  // namespace

hipError_t launch_fused_step(Storage st, const Fields& in, const Fields& out, const Scales& sc,
                             const Material& m, MatKind mk, const Grid& g, const Sources& s,
                             const Dft& d, int n, int run_step, hipStream_t stream) {
  // in-plane offsets are 32-bit byte offsets
  if ((unsigned long long)g.sxy * 8ull >= (1ull << 32)) return hipErrorInvalidValue;
  switch (st) {
    case kF32:
      if (mk == kMatFull) return launch_fused<float, kMatFull>(in, out, sc, m, g, s, d, n, run_step, stream);
      if (mk == kMatIdx) return launch_fused<float, kMatIdx>(in, out, sc, m, g, s, d, n, run_step, stream);
      return launch_fused<float, kMatIso>(in, out, sc, m, g, s, d, n, run_step, stream);
    case kF64:
      if (mk == kMatFull) return launch_fused<double, kMatFull>(in, out, sc, m, g, s, d, n, run_step, stream);
      if (mk == kMatIdx) return hipErrorInvalidValue;
      return launch_fused<double, kMatIso>(in, out, sc, m, g, s, d, n, run_step, stream);
    case kF16:
      if (mk == kMatFull) return hipErrorInvalidValue;
      if (mk == kMatIdx) return launch_fused<_Float16, kMatIdx>(in, out, sc, m, g, s, d, n, run_step, stream);
      return launch_fused<_Float16, kMatIso>(in, out, sc, m, g, s, d, n, run_step, stream);
    case kBF16:
      if (mk == kMatFull) return hipErrorInvalidValue;
      if (mk == kMatIdx) return launch_fused<__bf16, kMatIdx>(in, out, sc, m, g, s, d, n, run_step, stream);
      return launch_fused<__bf16, kMatIso>(in, out, sc, m, g, s, d, n, run_step, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace gd
