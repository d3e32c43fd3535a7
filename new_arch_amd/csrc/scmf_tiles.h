// scmf_tiles.h — spatial ordering of chains and the tiled NGP gridding.
//
// Global atomics run at ~27 G lane-ops/s on MI355X whatever their scope or
// address locality, i.e. ~10 ms for 2.7e8 beads.  The tiled gridding issues
// none on its common path:
//   1. chains are kept ordered by the spatial tile of their middle bead,
//      within each architecture block (so bead offsets stay a formula of the
//      slot index and the per-bead monomer types never move);
//   2. one workgroup per tile bins its chains' beads into an LDS window
//      (the tile plus a halo of h cells per side, uint16 counters packed in
//      pairs; with two monomer types the pair is a cell's two types), four
//      beads per lane from 16-byte loads, and stores the window: verbatim
//      for two types, else as 27 contiguous regions (low halo, interior,
//      high halo per axis);
//   3. one workgroup per tile sums, for each of its cells, the windows that
//      cover it (its own and up to seven neighbours).
// A bead outside its tile's window (a chain that drifted since the last
// ordering) is added with a global atomic instead, so results never depend
// on how well ordered the chains are; only speed does.
//
// Tiles split each axis into nt pieces of q or q+1 cells (the first `rem`
// pieces get q+1); a piece is never shorter than twice the halo, so every
// halo region of a window lies inside one neighbouring tile and every cell
// lies in the halo of at most one neighbour per axis.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "scmf_device.h"

namespace scmf {

struct TileGeo {
  int nt[3];          // tiles per axis
  int q[3], rem[3];   // tile i along an axis has q + (i < rem) cells
  int h[3];           // halo in cells (0 on an axis with a single tile)
  int E[3];           // largest tile edge, q + (rem > 0)
  int W[3];           // LDS window box edge, E + 2h
  int ntiles;         // nt0 nt1 nt2
  int win_cells;      // W0 W1 W2
  int win_words;      // T win_cells / 2, rounded up (uint16 pairs)
  int win_words_padded;  // win_words rounded up to a multiple of 4
  int slot_elems;     // uint16 elements of one window image in the image buffer
  int reg_off[27];    // element offset of region rx + 3 ry + 9 rz in an image
};

// Regions along an axis: 0 = low halo, 1 = interior, 2 = high halo.
__host__ __device__ __forceinline__ int tile_len(int q, int rem, int i) { return q + (i < rem ? 1 : 0); }
__host__ __device__ __forceinline__ int tile_first(int q, int rem, int i) {
  return i * q + (i < rem ? i : rem);
}
__host__ __device__ __forceinline__ int tile_index(int q, int rem, int c) {
  const int big = rem * (q + 1);
  return c < big ? c / (q + 1) : rem + (c - big) / q;
}

// floor(a / b) for 0 <= a, 0 < b, a + b < 2^21: (a + 1/2)/b lies at least
// 1/(2b) from an integer, more than the error of the fp32 product.
__device__ __forceinline__ int small_div(int a, float inv_b) {
  return (int)(((float)a + 0.5f) * inv_b);
}

// ---------------------------------------------------------------- ordering
// Key of every slot: architecture block, then tile of the chain's middle bead.
// Adds one to counter[key] for every active lane of the wave with one atomic
// per distinct key; returns the counter's value before the lane's own unit
// (the old value plus the lane's rank among the lanes with the same key).
// Slots in the previous order are mostly grouped by tile already, so a wave
// usually holds one or two keys and the tile counters see few atomics.
__device__ __forceinline__ uint32_t wave_add_one(uint32_t* counter, uint32_t key) {
  const int lane = (int)(threadIdx.x & 63u);
  unsigned long long todo = __ballot(1);
  uint32_t mine = 0u;
  while (todo) {
    const int leader = __ffsll(todo) - 1;
    const uint32_t k = __shfl(key, leader);
    const unsigned long long same = __ballot(key == k);
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(&counter[k], (uint32_t)__popcll(same));
    base = __shfl(base, leader);
    if (key == k) mine = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  return mine;
}

__global__ void __launch_bounds__(256) k_tile_keys(KParams kp, TileGeo g,
                                                   const float* __restrict__ x,
                                                   const float* __restrict__ y,
                                                   const float* __restrict__ z,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ count) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= kp.n_chains) return;
  const int p = arch_of_chain(kp, s);
  const int64_t a = kp.bead_base[p] + (s - kp.chain_base[p]) * kp.len[p] + kp.len[p] / 2;
  const int xi = fold((double)x[a], axis(kp, 0));
  const int yi = fold((double)y[a], axis(kp, 1));
  const int zi = fold((double)z[a], axis(kp, 2));
  const uint32_t t = (uint32_t)(tile_index(g.q[0], g.rem[0], xi) +
                                g.nt[0] * (tile_index(g.q[1], g.rem[1], yi) +
                                           g.nt[1] * tile_index(g.q[2], g.rem[2], zi)));
  const uint32_t key = (uint32_t)p * (uint32_t)g.ntiles + t;
  keys[s] = key;
  (void)wave_add_one(count, key);
}

// Exclusive scan of n counts into start[0..n] (one workgroup of 1024 lanes),
// clearing the fill counters.
__global__ void __launch_bounds__(1024) k_scan_counts(const uint32_t* __restrict__ count,
                                                      uint32_t* __restrict__ start,
                                                      uint32_t* __restrict__ fill, int n) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = min(n, b + per);
  uint32_t sum = 0;
  for (int i = b; i < e; ++i) sum += count[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    const uint32_t v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = t ? part[t - 1] : 0u;
  for (int i = b; i < e; ++i) {
    start[i] = run;
    run += count[i];
    fill[i] = 0u;
  }
  if (t == 1023) start[n] = part[1023];
}

// New slot of every old slot: its tile's first slot plus an arrival rank.
// (The order inside a tile is arbitrary; nothing computed depends on it.)
__global__ void __launch_bounds__(256) k_tile_scatter(int64_t n, const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ start,
                                                      uint32_t* __restrict__ fill,
                                                      uint32_t* __restrict__ new_slot) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const uint32_t k = keys[s];
  new_slot[s] = start[k] + wave_add_one(fill, k);
}

// Move every chain to its new slot (positions double-buffered, and the bead
// cells when they match the positions) and update the slot <-> chain maps.  One lane per bead; the slot of a bead follows
// from its architecture block by one fp64 product, corrected by one step.
__global__ void __launch_bounds__(256) k_permute(KParams kp, const float* __restrict__ x,
                                                 const float* __restrict__ y,
                                                 const float* __restrict__ z,
                                                 float* __restrict__ x2, float* __restrict__ y2,
                                                 float* __restrict__ z2,
                                                 const uint32_t* __restrict__ cells,
                                                 uint32_t* __restrict__ cells2,
                                                 const uint32_t* __restrict__ new_slot,
                                                 const uint32_t* __restrict__ slot_chain,
                                                 uint32_t* __restrict__ slot_chain2,
                                                 uint32_t* __restrict__ chain_slot) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t gb = (int64_t)blockIdx.x * 256 + threadIdx.x; gb < kp.n_beads; gb += stride) {
    const int p = arch_of_bead(kp, gb);
    const int len = kp.len[p];
    const int64_t rel = gb - kp.bead_base[p];
    int64_t q = (int64_t)((double)rel * kp.inv_len[p]);
    int64_t m = rel - q * len;
    if (m >= len) {
      m -= len;
      ++q;
    }
    if (m < 0) {
      m += len;
      --q;
    }
    const int64_t s = kp.chain_base[p] + q;
    const int64_t ns = new_slot[s];
    const int64_t dst = kp.bead_base[p] + (ns - kp.chain_base[p]) * len + m;
    x2[dst] = x[gb];
    y2[dst] = y[gb];
    z2[dst] = z[gb];
    if (cells) cells2[dst] = cells[gb];
    if (m == 0) {
      const uint32_t c = slot_chain[s];
      slot_chain2[ns] = c;
      chain_slot[c] = (uint32_t)ns;
    }
  }
}

// ---------------------------------------------------------------- gridding
// Offset of cell c from a window origin in [-h, n), folded into [0, n).
__device__ __forceinline__ int wrap_rel(int c, int origin, int n) {
  int d = c - origin;
  if (d < 0) d += n;
  if (d >= n) d -= n;
  return d;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32v4 __attribute__((ext_vector_type(4)));

// Packed cell of a bead, as the sweep stores it: 10 bits per axis.
constexpr int CELL_AXIS_MAX = 1024;
__device__ __forceinline__ uint32_t pack_cell(int xi, int yi, int zi) {
  return (uint32_t)xi | ((uint32_t)yi << 10) | ((uint32_t)zi << 20);
}

constexpr int MONO_LDS = 4096;     // architectures up to this length read types from LDS
constexpr uint32_t OUT_OF_WINDOW = 1u << 30;  // the sum of three table entries stays below 2^32
constexpr int GRID_THREADS = 1024;
// A pass of the binning loop: every lane bins four beads.
constexpr int PASS_BEADS = 4 * GRID_THREADS;
// Beads binned between two normalisations of the window (whole passes, at
// most 32767): each uint16 half then grows by at most this much from below
// 2^15, so it cannot wrap.
constexpr int NORM_BEADS = (32767 / PASS_BEADS) * PASS_BEADS;
static_assert(NORM_BEADS > 0 && NORM_BEADS <= 32767, "normalisation interval");

// The general binning of three coordinates (the tiled gridding needs it only
// for beads outside the box or off power-of-two cells).
__device__ __forceinline__ void fold3_general(float x, float y, float z, const Axis ax,
                                           const Axis ay, const Axis az, int* xi, int* yi,
                                           int* zi) {
  *xi = fold((double)x, ax);
  *yi = fold((double)y, ay);
  *zi = fold((double)z, az);
}

// Counts `v` more beads of global cell `cell` with a global atomic.
template <int BITS>
__device__ __forceinline__ void add_global(SCMF_GLOBAL uint32_t* counts, uint32_t* overflow,
                                           uint32_t cell, uint32_t v) {
  if (BITS == 32) {
    atomicAdd((uint32_t*)&counts[cell], v);
  } else {
    const uint32_t sh = (cell & 1u) * 16u;
    const uint32_t old = atomicAdd((uint32_t*)&counts[cell >> 1], v << sh);
    if (((old >> sh) & 0xFFFFu) + v > 0xFFFFu) atomicOr(overflow, 1u);
  }
}

// Move every uint16 half of the window that has reached 2^15 to the global
// counts and clear it (only tiles with more than 65535 beads need this).
template <int BITS, bool PAIR>
__device__ void normalise_window(uint32_t* win, const TileGeo& g, KPC& kp, int ox, int oy, int oz,
                                 SCMF_GLOBAL uint32_t* counts, uint32_t* overflow) {
  // four words per lane and test (the window is padded to a multiple of four)
  for (int w4 = threadIdx.x; 4 * w4 < g.win_words; w4 += GRID_THREADS) {
    const uint4 q = ((const uint4*)win)[w4];
    if (!((q.x | q.y | q.z | q.w) & 0x80008000u)) continue;
    for (int j = 0; j < 4; ++j) {
    const int w = 4 * w4 + j;
    const uint32_t v = win[w];
    if (!(v & 0x80008000u)) continue;
    uint32_t keep = v;
    for (int half = 0; half < 2; ++half) {
      const uint32_t val = (v >> (16 * half)) & 0xFFFFu;
      if (val < 0x8000u) continue;
      keep &= ~(0xFFFFu << (16 * half));
      // window cell and type of the half (see window_half)
      const int wc = 2 * w + half;
      const int type = PAIR ? half : wc / g.win_cells;
      const int r = PAIR ? w : wc - type * g.win_cells;
      const int wx = r % g.W[0], wy = (r / g.W[0]) % g.W[1], wz = r / (g.W[0] * g.W[1]);
      const int cx = wrap_rel(ox + wx, 0, kp.nx), cy = wrap_rel(oy + wy, 0, kp.ny),
                cz = wrap_rel(oz + wz, 0, kp.nz);
      add_global<BITS>(counts, overflow,
                       (uint32_t)type * kp.n_cells + (uint32_t)(cx + kp.nx * (cy + kp.ny * cz)), val);
    }
    win[w] = keep;
    }
  }
}

// One workgroup per tile: bin the beads of the tile's chains into the LDS
// window, then store the window (verbatim for two monomer types, else as 27
// regions).  Beads outside the window go straight to the final counts with a
// global atomic.  Positions are read four beads per lane (16-byte loads; the
// arrays are padded to a multiple of four beads), later passes' loads in
// flight while the current ones are binned.
// With CELLS the cells the last sweep stored for the same positions are read
// instead (4 bytes a bead, no binning).
//
// Window layout: element (type, window cell c) is uint16 half number
// type * win_cells + c of the window; with PAIR (two monomer types) it is
// half `type` of word c instead, so a bead's counter word is addressed by
// its cell alone (the axis tables then hold byte offsets) and its type only
// selects the increment, 1 or 1 << 16.
template <bool PAIR>
__device__ __forceinline__ uint32_t window_half(const uint32_t* win, uint32_t wc, uint32_t wcl) {
  if (PAIR) {
    const uint32_t type = wc >= wcl ? 1u : 0u;
    return (win[wc - (type ? wcl : 0u)] >> (type << 4)) & 0xFFFFu;
  }
  return (win[wc >> 1] >> ((wc & 1u) << 4)) & 0xFFFFu;
}

template <int BITS, bool POW2, bool CELLS, bool PAIR>
__global__ void __launch_bounds__(GRID_THREADS) k_grid_tiles(const KParams* __restrict__ Ps, int sys,
                                                             TileGeo g,
                                                             const uint32_t* __restrict__ tile_start,
                                                             uint16_t* __restrict__ images,
                                                             unsigned long long* __restrict__ fallback) {
  extern __shared__ __align__(16) uint32_t win[];
  // types of an architecture's monomers (with PAIR, 16 x type: the shift of
  // the increment), the first three repeated after the last so that the four
  // beads of a lane read consecutive entries without wrapping
  __shared__ uint8_t mono_s[MONO_LDS + 4];
  __shared__ int reg_s[27];
  KPC& kp = param(Ps, (uint32_t)sys);
  if (step_skipped(kp)) return;
  if (threadIdx.x < 27) reg_s[threadIdx.x] = g.reg_off[threadIdx.x];
  // per-axis tables after the window: cell -> offset in the window (scaled
  // by the strides of the window box; with PAIR in bytes), or OUT_OF_WINDOW
  uint32_t* const tab_x = win + g.win_words_padded;
  uint32_t* const tab_y = tab_x + kp.nx;
  uint32_t* const tab_z = tab_y + kp.ny;
  const SCMF_GLOBAL float* __restrict__ x = gptr((const float*)kp.x);
  const SCMF_GLOBAL float* __restrict__ y = gptr((const float*)kp.y);
  const SCMF_GLOBAL float* __restrict__ z = gptr((const float*)kp.z);
  const SCMF_GLOBAL uint32_t* __restrict__ cellw = gptr((const uint32_t*)kp.cells);
  const SCMF_GLOBAL uint8_t* __restrict__ btype = gptr((const uint8_t*)kp.btype);
  SCMF_GLOBAL uint32_t* __restrict__ counts = gptr((uint32_t*)kp.counts);
  uint32_t* __restrict__ overflow = (uint32_t*)kp.overflow;
  const int t = blockIdx.x;
  const int tx = t % g.nt[0], ty = (t / g.nt[0]) % g.nt[1], tz = t / (g.nt[0] * g.nt[1]);
  const int sx = tile_len(g.q[0], g.rem[0], tx), sy = tile_len(g.q[1], g.rem[1], ty),
            sz = tile_len(g.q[2], g.rem[2], tz);
  const int ox = tile_first(g.q[0], g.rem[0], tx) - g.h[0],
            oy = tile_first(g.q[1], g.rem[1], ty) - g.h[1],
            oz = tile_first(g.q[2], g.rem[2], tz) - g.h[2];
  const int wxm = sx + 2 * g.h[0], wym = sy + 2 * g.h[1], wzm = sz + 2 * g.h[2];
  for (int w = threadIdx.x; w < g.win_words_padded; w += GRID_THREADS) win[w] = 0u;
  for (int c = threadIdx.x; c < kp.nx; c += GRID_THREADS) {
    const int d = wrap_rel(c, ox, kp.nx);
    tab_x[c] = d < wxm ? (uint32_t)d * (PAIR ? 4u : 1u) : OUT_OF_WINDOW;
  }
  for (int c = threadIdx.x; c < kp.ny; c += GRID_THREADS) {
    const int d = wrap_rel(c, oy, kp.ny);
    tab_y[c] = d < wym ? (uint32_t)(d * g.W[0]) * (PAIR ? 4u : 1u) : OUT_OF_WINDOW;
  }
  for (int c = threadIdx.x; c < kp.nz; c += GRID_THREADS) {
    const int d = wrap_rel(c, oz, kp.nz);
    tab_z[c] = d < wzm ? (uint32_t)(d * g.W[0] * g.W[1]) * (PAIR ? 4u : 1u) : OUT_OF_WINDOW;
  }
  int64_t pop = 0;
  for (int p = 0; p < kp.P; ++p)
    pop += (int64_t)(tile_start[p * g.ntiles + t + 1] - tile_start[p * g.ntiles + t]) * kp.len[p];
  // A tile of at most 65,535 beads cannot wrap a 16-bit window counter.  A
  // larger tile is binned optimistically first, and its window checked
  // afterwards: every increment adds 1 to one uint16 half, a wrap of a half
  // lowers the sum of all halves by at least 65535 (a carry out of a low half
  // takes 65536 from it and adds 1 to the high half; a carry out of a word is
  // lost), and nothing else lowers it, so the window is exact exactly when
  // its halves sum to the number of beads counted in it (below 2^32).
  // Otherwise the window is cleared and the tile binned again, normalising
  // every NORM_BEADS beads so no half can wrap, with the beads outside the
  // window (already counted globally) skipped.  (The check replaced atomics
  // that returned the old words to be watched for bit 15: the returns made
  // every pass wait for its atomics.)
  const bool checked = pop > 65535;
  __shared__ uint32_t redo_s;
  __shared__ uint32_t spare_s;  // target of the atomics of beads not counted in the window
  __shared__ uint32_t missed_s;  // beads of the tile counted with global atomics
  __shared__ uint32_t here_s;    // beads counted in the window (checked tiles)
  __shared__ uint32_t wsum_s;    // sum of the window's uint16 halves (checked tiles)
  if (threadIdx.x == 0) {
    redo_s = 0u;
    spare_s = 0u;
    missed_s = 0u;
    here_s = 0u;
    wsum_s = 0u;
  }
  const Axis ax0 = axis(kp, 0), ax1 = axis(kp, 1), ax2 = axis(kp, 2);
  // (kept in registers: the loop's global atomics would make the compiler
  // reload them from the parameter block)
  const uint32_t n_cells = (uint32_t)kp.n_cells, nx = (uint32_t)kp.nx, ny = (uint32_t)kp.ny;
  uint32_t missed = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
  const bool safe = attempt == 1;        // the normalising pass
  const bool track = checked && !safe;   // the window checked for wraps
  if (safe) {
    __syncthreads();
    if (redo_s == 0u) break;
    for (int w = threadIdx.x; w < g.win_words_padded; w += GRID_THREADS) win[w] = 0u;
  }
  uint32_t nhere = 0;
  int since = 0;
  for (int p = 0; p < kp.P; ++p) {
    const int64_t s0 = tile_start[p * g.ntiles + t], s1 = tile_start[p * g.ntiles + t + 1];
    if (s1 <= s0) continue;
    const int len = kp.len[p];
    const int64_t bb = kp.bead_base[p];
    const int64_t b0 = bb + (s0 - kp.chain_base[p]) * len;
    const int64_t b1 = bb + (s1 - kp.chain_base[p]) * len;
    const bool tab = len <= MONO_LDS;
    __syncthreads();
    if (tab)
      for (int i = threadIdx.x; i < len + 3; i += GRID_THREADS)
        mono_s[i] = (uint8_t)(kp.monotype[kp.mono_off[p] + i % len] * (PAIR ? 16 : 1));
    __syncthreads();
    const double inv_len = kp.inv_len[p];
    // The type source is fixed per architecture; two copies of the loop keep
    // a global type load from sharing registers with the LDS one (which
    // would make every LDS read wait for the prefetched positions).
    // With CTY the lane's four monomer types are read once: the chain length
    // divides the pass, so the lane's monomer indices never change.
    auto bin_range = [&](auto tab_c, auto cty_c) {
      constexpr bool TAB = decltype(tab_c)::value;
      constexpr bool CTY = decltype(cty_c)::value;
      // Each lane takes four beads per pass and keeps the loads of the next
      // PF passes in flight while it bins these (one pass a lane is ~4 MB in
      // flight on the whole GPU, ~16 KB per CU).  Loads are issued
      // unconditionally (lanes past the range re-read the last group
      // of the range), so the wait before binning counts only the current
      // group and the later ones stay in flight.
      constexpr int PF = CELLS ? 4 : 2;
      const int64_t last = (b1 - 1) & ~(int64_t)3;
      int64_t gq = (b0 & ~(int64_t)3) + 4 * (int64_t)threadIdx.x;
      // monomer index of the lane's first bead (floor mod len), then advanced
      const int step_m = PASS_BEADS % len;
      int m0;
      {
        const int64_t rel = gq - bb;
        const int64_t qd = (int64_t)__builtin_floor((double)rel * inv_len);
        m0 = (int)(rel - qd * len);
        if (m0 < 0) m0 += len;
        if (m0 >= len) m0 -= len;
      }
      uint32_t tyc[4] = {};
      if constexpr (CTY) {
#pragma unroll
        for (int e = 0; e < 4; ++e) tyc[e] = (uint32_t)mono_s[m0 + e];
      }
      auto load = [&](int64_t gl, f32x4& Xl, f32x4& Yl, f32x4& Zl, u32v4& Ql) {
        const int64_t a = gl < b1 ? gl : last;
        if constexpr (CELLS) {
          Ql = *(const SCMF_GLOBAL u32v4*)(cellw + a);
        } else {
          Xl = *(const SCMF_GLOBAL f32x4*)(x + a);
          Yl = *(const SCMF_GLOBAL f32x4*)(y + a);
          Zl = *(const SCMF_GLOBAL f32x4*)(z + a);
        }
      };
      f32x4 Xr[PF] = {}, Yr[PF] = {}, Zr[PF] = {};
      u32v4 Qr[PF] = {};
      // (issued in pass order: loads complete in order, and the wait for
      // pass k must not cover the later ones)
#pragma unroll
      for (int k = 0; k < PF; ++k) {
        load(gq + (int64_t)k * PASS_BEADS, Xr[k], Yr[k], Zr[k], Qr[k]);
        __builtin_amdgcn_sched_barrier(0);
      }
      // PF passes per iteration: slot k of the ring always holds pass k
      // (mod PF), so no register of a load in flight is ever copied (a copy
      // would wait for the load).
      int64_t base = b0 & ~(int64_t)3;
      while (base < b1) {
#pragma unroll
      for (int k = 0; k < PF; ++k) {
      if (base < b1) {
        const f32x4 X = Xr[k], Y = Yr[k], Z = Zr[k];
        const u32v4 Q = Qr[k];
        load(gq + (int64_t)PF * PASS_BEADS, Xr[k], Yr[k], Zr[k], Qr[k]);
        if (gq < b1) {
          const float px[4] = {X.x, X.y, X.z, X.w}, py[4] = {Y.x, Y.y, Y.z, Y.w},
                      pz[4] = {Z.x, Z.y, Z.z, Z.w};
          const uint32_t pq[4] = {Q.x, Q.y, Q.z, Q.w};
          // beads e in [elo, ehi) of the four belong to the range
          const int elo = gq < b0 ? (int)(b0 - gq) : 0;
          const int ehi = b1 - gq < 4 ? (int)(b1 - gq) : 4;
          // cell of bead e.  A bead outside the range decodes to a cell of
          // the box too (so its table reads stay in bounds): its cell word is
          // another bead's or the padding's, which is zeroed when allocated,
          // and the fold maps any coordinate, even a non-finite one, into
          // the box.
          auto decode = [&](int e, int* xi, int* yi, int* zi) {
            if constexpr (CELLS) {
              *xi = (int)(pq[e] & 1023u);
              *yi = (int)((pq[e] >> 10) & 1023u);
              *zi = (int)(pq[e] >> 20);
            } else {
              const bool quick = POW2 && (fold_pow2(px[e], ax0, xi) & fold_pow2(py[e], ax1, yi) &
                                          fold_pow2(pz[e], ax2, zi));
              if (!quick) fold3_general(px[e], py[e], pz[e], ax0, ax1, ax2, xi, yi, zi);
            }
          };
          // The LDS work of the four beads is issued in two batches (the
          // table and type reads, then the counter atomics), so each lane
          // waits for LDS twice per pass rather than twice per bead.
          uint32_t sp[4], ty[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            int xi, yi, zi;
            decode(e, &xi, &yi, &zi);
            sp[e] = tab_x[xi] + tab_y[yi] + tab_z[zi];
            if (CTY)
              ty[e] = tyc[e];
            else if (TAB)
              ty[e] = (uint32_t)mono_s[m0 + e];
            else
              ty[e] = e >= elo && e < ehi ? (uint32_t)btype[gq + e] * (PAIR ? 16u : 1u) : 0u;
          }
          // (keeps the compiler from using each bead's reads before the
          // next bead's are issued)
          __builtin_amdgcn_sched_barrier(0);
          // Every lane issues all four atomics: a bead that is not counted
          // here adds 0 to a spare word, so no branch (and no wait the
          // compiler would place before one) separates them.
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const bool here = e >= elo && e < ehi && sp[e] < OUT_OF_WINDOW;
            uint32_t* word;
            uint32_t inc;
            if (PAIR) {
              word = (uint32_t*)((char*)win + sp[e]);
              inc = 1u << ty[e];
            } else {
              const uint32_t wc = __umul24(ty[e], (uint32_t)g.win_cells) + sp[e];
              word = &win[wc >> 1];
              inc = 1u << ((wc & 1u) << 4);
            }
            atomicAdd(here ? word : &spare_s, here ? inc : 0u);
            nhere += here ? 1u : 0u;
          }
          // beads outside the window (one wave-uniform test per pass)
          bool out[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) out[e] = e >= elo && e < ehi && sp[e] >= OUT_OF_WINDOW;
          if (!safe && __ballot(out[0] | out[1] | out[2] | out[3])) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              if (out[e]) {
                int xi, yi, zi;
                decode(e, &xi, &yi, &zi);
                ++missed;
                const uint32_t type = PAIR ? ty[e] >> 4 : ty[e];
                add_global<BITS>(counts, overflow,
                                 type * n_cells + (uint32_t)(xi + nx * (yi + ny * zi)), 1u);
              }
            }
          }
        }
        gq += PASS_BEADS;
        m0 += step_m;
        if (m0 >= len) m0 -= len;
        if (safe && (since += PASS_BEADS) >= NORM_BEADS) {
          since = 0;
          __syncthreads();
          normalise_window<BITS, PAIR>(win, g, kp, ox, oy, oz, counts, overflow);
          __syncthreads();
        }
        base += PASS_BEADS;
      }
      }
      }
    };
    if (tab && PASS_BEADS % len == 0)
      bin_range(std::true_type{}, std::true_type{});
    else if (tab)
      bin_range(std::true_type{}, std::false_type{});
    else
      bin_range(std::false_type{}, std::false_type{});
  }
  if (track) {
    for (int off = 32; off > 0; off >>= 1) nhere += __shfl_xor(nhere, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&here_s, nhere);
    __syncthreads();
    uint32_t sum = 0;
    for (int w4 = threadIdx.x; 4 * w4 < g.win_words_padded; w4 += GRID_THREADS) {
      const uint4 q = ((const uint4*)win)[w4];
      sum += ((q.x & 0xFFFFu) + (q.x >> 16)) + ((q.y & 0xFFFFu) + (q.y >> 16)) +
             ((q.z & 0xFFFFu) + (q.z >> 16)) + ((q.w & 0xFFFFu) + (q.w >> 16));
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&wsum_s, sum);
    __syncthreads();
    if (threadIdx.x == 0 && wsum_s != here_s) redo_s = 1u;
  }
  }
  // beads counted globally: summed in LDS, one global atomic per tile (one
  // per wave on the single counter contended when tiles end together)
  for (int off = 32; off > 0; off >>= 1) missed += __shfl_xor(missed, off);
  if ((threadIdx.x & 63) == 0 && missed) atomicAdd(&missed_s, missed);
  __syncthreads();
  if (threadIdx.x == 0 && missed_s) atomicAdd(fallback, (unsigned long long)missed_s);
  if (PAIR) {
    // the window verbatim, 16 bytes a lane (images are 128-byte aligned)
    uint4* __restrict__ dst = (uint4*)(images + (int64_t)t * g.slot_elems);
    for (int w4 = threadIdx.x; 4 * w4 < g.win_words_padded; w4 += GRID_THREADS)
      dst[w4] = ((const uint4*)win)[w4];
    return;
  }
  // store the window as 27 regions
  uint16_t* __restrict__ img = images + (int64_t)t * g.slot_elems;
  const int sz3[3] = {sx, sy, sz};
  for (int r = 0; r < 27; ++r) {
    const int rx = r % 3, ry = (r / 3) % 3, rz = r / 9;
    const int ex = rx == 1 ? sz3[0] : g.h[0], ey = ry == 1 ? sz3[1] : g.h[1],
              ez = rz == 1 ? sz3[2] : g.h[2];
    const int fx = rx == 0 ? 0 : (rx == 1 ? g.h[0] : g.h[0] + sx);
    const int fy = ry == 0 ? 0 : (ry == 1 ? g.h[1] : g.h[1] + sy);
    const int fz = rz == 0 ? 0 : (rz == 1 ? g.h[2] : g.h[2] + sz);
    const int nrow = ex * ey * ez;
    const int n_r = kp.T * nrow;
    if (n_r == 0) continue;
    const float iex = 1.0f / (float)ex, iey = 1.0f / (float)ey, irow = 1.0f / (float)nrow;
    // Two elements per lane, stored as one word: a region starts at an even
    // element of a 4-byte aligned image and is allocated with an even count
    // (an odd region's last word carries a spare element nothing reads).
    uint32_t* __restrict__ dst = (uint32_t*)(img + reg_s[r]);
    // (every product below is of values under 2^24)
    const uint32_t W0 = (uint32_t)g.W[0], W01 = (uint32_t)(g.W[0] * g.W[1]), wcl = (uint32_t)g.win_cells;
    const uint32_t base_off = (uint32_t)fx + W0 * (uint32_t)fy + W01 * (uint32_t)fz;
    // window-cell steps from the last cell of a row to the first of the next
    // row, plane and type
    const uint32_t row_step = W0 - (uint32_t)(ex - 1);
    const uint32_t plane_step = W01 - (uint32_t)(ex - 1) - W0 * (uint32_t)(ey - 1);
    const uint32_t type_step = wcl - (uint32_t)(ex - 1) - W0 * (uint32_t)(ey - 1) -
                               W01 * (uint32_t)(ez - 1);
    const int n_pairs = (n_r + 1) >> 1;
    for (int j = threadIdx.x; j < n_pairs; j += GRID_THREADS) {
      const int e = 2 * j;
      const int type = small_div(e, irow);
      const int l = e - (int)__umul24((uint32_t)type, (uint32_t)nrow);
      const int lyz = small_div(l, iex);
      const int lx = l - (int)__umul24((uint32_t)lyz, (uint32_t)ex);
      const int lz = small_div(lyz, iey);
      const int ly = lyz - (int)__umul24((uint32_t)lz, (uint32_t)ey);
      const uint32_t wc0 = __umul24((uint32_t)type, wcl) + base_off + (uint32_t)lx +
                           __umul24(W0, (uint32_t)ly) + __umul24(W01, (uint32_t)lz);
      const uint32_t wc1 = wc0 + (lx + 1 < ex   ? 1u
                                  : ly + 1 < ey ? row_step
                                  : lz + 1 < ez ? plane_step
                                  : e + 1 < n_r ? type_step
                                                : 0u);
      const uint32_t v0 = window_half<PAIR>(win, wc0, wcl);
      const uint32_t v1 = window_half<PAIR>(win, wc1, wcl);
      dst[j] = v0 | (v1 << 16);
    }
  }
}

// One workgroup per tile: every cell of the tile gets its window's interior
// value, the halo values of the neighbouring windows that reach it, and the
// globally counted beads already in the counts.
constexpr int SUM_THREADS = 1024;

// Workgroups are dealt to the 8 XCDs in turn; give each XCD a contiguous run
// of tiles so that neighbouring tiles (which share cache lines of the counts)
// meet in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int x = b & 7, i = b >> 3, per = n >> 3, extra = n & 7;
  return x * per + (x < extra ? x : extra) + i;
}

template <int BITS>
__global__ void __launch_bounds__(SUM_THREADS) k_tiles_sum(const KParams* __restrict__ Ps, int sys,
                                                   TileGeo g,
                                                   const uint16_t* __restrict__ images,
                                                   int advance) {
  KPC& kp = param(Ps, (uint32_t)sys);
  if (step_skipped(kp)) return;
  if (advance) advance_sweep_word(kp);
  SCMF_GLOBAL uint32_t* __restrict__ counts = gptr((uint32_t*)kp.counts);
  SCMF_GLOBAL uint16_t* __restrict__ counts16 = gptr((uint16_t*)kp.counts);
  uint32_t* __restrict__ overflow = (uint32_t*)kp.overflow;
  __shared__ int reg_s[27];
  if (threadIdx.x < 27) reg_s[threadIdx.x] = g.reg_off[threadIdx.x];
  __syncthreads();
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int v0 = t % g.nt[0], v1 = (t / g.nt[0]) % g.nt[1], v2 = t / (g.nt[0] * g.nt[1]);
  const int s0 = tile_len(g.q[0], g.rem[0], v0), s1 = tile_len(g.q[1], g.rem[1], v1),
            s2 = tile_len(g.q[2], g.rem[2], v2);
  const int f0 = tile_first(g.q[0], g.rem[0], v0), f1 = tile_first(g.q[1], g.rem[1], v1),
            f2 = tile_first(g.q[2], g.rem[2], v2);
  // neighbouring windows along each axis (periodic)
  const int d0 = v0 == 0 ? g.nt[0] - 1 : v0 - 1, u0 = v0 == g.nt[0] - 1 ? 0 : v0 + 1;
  const int d1 = v1 == 0 ? g.nt[1] - 1 : v1 - 1, u1 = v1 == g.nt[1] - 1 ? 0 : v1 + 1;
  const int d2 = v2 == 0 ? g.nt[2] - 1 : v2 - 1, u2 = v2 == g.nt[2] - 1 ? 0 : v2 + 1;
  const int ncell = s0 * s1 * s2;
  const int n_int = kp.T * ncell;
  const float icell = 1.0f / (float)ncell, isx = 1.0f / (float)s0, isy = 1.0f / (float)s1;
  const int64_t zero_idx = (int64_t)g.ntiles * g.slot_elems;  // an element that stays 0
  for (int e = threadIdx.x; e < n_int; e += SUM_THREADS) {
    const int type = small_div(e, icell);
    const int l = e - type * ncell;
    const int lyz = small_div(l, isx);
    const int i0 = l - lyz * s0;
    const int i2 = small_div(lyz, isy);
    const int i1 = lyz - i2 * s1;
    // per axis: the own window's interior, and at most one neighbouring
    // window whose halo covers the cell (the window below if the cell is in
    // the low band, above if in the high band)
    const bool lo0 = i0 < g.h[0], hi0 = i0 >= s0 - g.h[0];
    const bool lo1 = i1 < g.h[1], hi1 = i1 >= s1 - g.h[1];
    const bool lo2 = i2 < g.h[2], hi2 = i2 >= s2 - g.h[2];
    const int wn0 = lo0 ? d0 : u0, rn0 = lo0 ? 2 : 0, ln0 = lo0 ? i0 : i0 - (s0 - g.h[0]);
    const int wn1 = lo1 ? d1 : u1, rn1 = lo1 ? 2 : 0, ln1 = lo1 ? i1 : i1 - (s1 - g.h[1]);
    const int wn2 = lo2 ? d2 : u2, rn2 = lo2 ? 2 : 0, ln2 = lo2 ? i2 : i2 - (s2 - g.h[2]);
    const bool has0 = lo0 || hi0, has1 = lo1 || hi1, has2 = lo2 || hi2;
    const uint32_t cell = (uint32_t)type * kp.n_cells +
                          (uint32_t)((f0 + i0) + kp.nx * ((f1 + i1) + kp.ny * (f2 + i2)));
    int64_t idx[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const bool a = c & 1, b = c & 2, z = c & 4;
      const int w0 = a ? wn0 : v0, r0 = a ? rn0 : 1, l0 = a ? ln0 : i0, e0 = a ? g.h[0] : s0;
      const int w1 = b ? wn1 : v1, r1 = b ? rn1 : 1, l1 = b ? ln1 : i1, e1 = b ? g.h[1] : s1;
      const int w2 = z ? wn2 : v2, r2 = z ? rn2 : 1, l2 = z ? ln2 : i2, e2 = z ? g.h[2] : s2;
      const bool live = (!a || has0) && (!b || has1) && (!z || has2);
      const int u = w0 + g.nt[0] * (w1 + g.nt[1] * w2);
      idx[c] = live ? (int64_t)u * g.slot_elems + reg_s[r0 + 3 * (r1 + 3 * r2)] +
                          ((type * e2 + l2) * e1 + l1) * e0 + l0
                    : zero_idx;
    }
    uint32_t sum = BITS == 32 ? counts[cell] : (uint32_t)counts16[cell];
#pragma unroll
    for (int c = 0; c < 8; ++c) sum += images[idx[c]];
    if (BITS == 32) {
      counts[cell] = sum;
    } else {
      if (sum > 0xFFFFu) {
        atomicOr(overflow, 1u);
        sum = 0xFFFFu;
      }
      counts16[cell] = (uint16_t)sum;
    }
  }
}

// Two monomer types: the images are the windows themselves (word c holds
// both types' counters of window cell c).  One lane per cell: the tile's own
// window and, on each axis where the cell lies in a band of width h, the
// neighbouring window whose halo covers it; both types summed at once.
template <int BITS>
__global__ void __launch_bounds__(SUM_THREADS) k_tiles_sum_pair(const KParams* __restrict__ Ps,
                                                                int sys, TileGeo g,
                                                                const uint32_t* __restrict__ img,
                                                                int advance) {
  KPC& kp = param(Ps, (uint32_t)sys);
  if (step_skipped(kp)) return;
  if (advance) advance_sweep_word(kp);
  SCMF_GLOBAL uint32_t* __restrict__ counts = gptr((uint32_t*)kp.counts);
  SCMF_GLOBAL uint16_t* __restrict__ counts16 = gptr((uint16_t*)kp.counts);
  uint32_t* __restrict__ overflow = (uint32_t*)kp.overflow;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int v0 = t % g.nt[0], v1 = (t / g.nt[0]) % g.nt[1], v2 = t / (g.nt[0] * g.nt[1]);
  const int s0 = tile_len(g.q[0], g.rem[0], v0), s1 = tile_len(g.q[1], g.rem[1], v1),
            s2 = tile_len(g.q[2], g.rem[2], v2);
  const int f0 = tile_first(g.q[0], g.rem[0], v0), f1 = tile_first(g.q[1], g.rem[1], v1),
            f2 = tile_first(g.q[2], g.rem[2], v2);
  // neighbouring tiles along each axis (periodic) and the length of the one below
  const int d0 = v0 == 0 ? g.nt[0] - 1 : v0 - 1, u0 = v0 == g.nt[0] - 1 ? 0 : v0 + 1;
  const int d1 = v1 == 0 ? g.nt[1] - 1 : v1 - 1, u1 = v1 == g.nt[1] - 1 ? 0 : v1 + 1;
  const int d2 = v2 == 0 ? g.nt[2] - 1 : v2 - 1, u2 = v2 == g.nt[2] - 1 ? 0 : v2 + 1;
  const int sd0 = tile_len(g.q[0], g.rem[0], d0), sd1 = tile_len(g.q[1], g.rem[1], d1),
            sd2 = tile_len(g.q[2], g.rem[2], d2);
  const int64_t words = g.slot_elems / 2;  // 32-bit words of one image
  const int64_t zero_word = (int64_t)g.ntiles * words;
  const uint32_t W0 = (uint32_t)g.W[0], W01 = (uint32_t)(g.W[0] * g.W[1]);
  const int ncell = s0 * s1 * s2;
  const float isx = 1.0f / (float)s0, isy = 1.0f / (float)s1;
  for (int e = threadIdx.x; e < ncell; e += SUM_THREADS) {
    const int lyz = small_div(e, isx);
    const int i0 = e - lyz * s0;
    const int i2 = small_div(lyz, isy);
    const int i1 = lyz - i2 * s1;
    // per axis: the cell's coordinate in the own window, and (if the cell is
    // in the low or high band) the neighbour and the coordinate in its window
    const bool lo0 = i0 < g.h[0], hi0 = i0 >= s0 - g.h[0];
    const bool lo1 = i1 < g.h[1], hi1 = i1 >= s1 - g.h[1];
    const bool lo2 = i2 < g.h[2], hi2 = i2 >= s2 - g.h[2];
    const int n0 = lo0 ? d0 : u0, c0n = lo0 ? sd0 + g.h[0] + i0 : i0 - (s0 - g.h[0]);
    const int n1 = lo1 ? d1 : u1, c1n = lo1 ? sd1 + g.h[1] + i1 : i1 - (s1 - g.h[1]);
    const int n2 = lo2 ? d2 : u2, c2n = lo2 ? sd2 + g.h[2] + i2 : i2 - (s2 - g.h[2]);
    const bool has0 = lo0 || hi0, has1 = lo1 || hi1, has2 = lo2 || hi2;
    const uint32_t cell = (uint32_t)((f0 + i0) + kp.nx * ((f1 + i1) + kp.ny * (f2 + i2)));
    // all eight loads issued at once; an absent contribution reads the
    // zeroed word after the last image
    uint32_t wv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const bool a = c & 1, b = c & 2, z = c & 4;
      const bool live = (!a || has0) && (!b || has1) && (!z || has2);
      const int w0 = a ? n0 : v0, w1 = b ? n1 : v1, w2 = z ? n2 : v2;
      const uint32_t x0 = (uint32_t)(a ? c0n : g.h[0] + i0);
      const uint32_t x1 = (uint32_t)(b ? c1n : g.h[1] + i1);
      const uint32_t x2 = (uint32_t)(z ? c2n : g.h[2] + i2);
      const int64_t u = w0 + g.nt[0] * (w1 + g.nt[1] * w2);
      wv[c] = img[live ? u * words + x0 + W0 * x1 + W01 * x2 : zero_word];
    }
    uint32_t sum0 = BITS == 32 ? counts[cell] : (uint32_t)counts16[cell];
    uint32_t sum1 = BITS == 32 ? counts[kp.n_cells + cell] : (uint32_t)counts16[kp.n_cells + cell];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      sum0 += wv[c] & 0xFFFFu;
      sum1 += wv[c] >> 16;
    }
    if (BITS == 32) {
      counts[cell] = sum0;
      counts[kp.n_cells + cell] = sum1;
    } else {
      if ((sum0 | sum1) > 0xFFFFu) {
        atomicOr(overflow, 1u);
        sum0 = min(sum0, 0xFFFFu);
        sum1 = min(sum1, 0xFFFFu);
      }
      counts16[cell] = (uint16_t)sum0;
      counts16[kp.n_cells + cell] = (uint16_t)sum1;
    }
  }
}

}  // namespace scmf
