// scmf_tiles.h — spatial ordering of chains and the tiled NGP gridding.
//
// Global atomics run at ~27 G lane-ops/s on MI355X whatever their scope or
// address locality, i.e. ~10 ms for 2.7e8 beads.  The tiled gridding issues
// none on its common path:
//   1. chains are kept ordered by the spatial tile of their middle bead,
//      within each architecture block (so bead offsets stay a formula of the
//      slot index and the per-bead monomer types never move);
//   2. one workgroup per tile bins its chains' beads into an LDS window
//      (tile + halo, uint16 counters packed in pairs) and stores the window
//      to a scratch buffer with plain coalesced stores;
//   3. one lane per pair of cells sums the windows that cover it.
// A bead outside its tile's window (a chain that drifted since the last
// ordering) is added with a global atomic instead, so results never depend
// on how well ordered the chains are; only speed does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scmf_device.h"

namespace scmf {

struct TileGeo {
  int ts[3];        // tile edge in cells
  int nt[3];        // tiles per axis
  int h;            // halo in cells
  int W[3];         // window edge = ts + 2h (<= grid edge)
  int ntiles;       // nt0 nt1 nt2
  int win_cells;    // W0 W1 W2
  int win_words;    // T win_cells / 2, rounded up (uint16 pairs)
};

// ---------------------------------------------------------------- ordering
// Key of every slot: architecture block, then tile of the chain's middle bead.
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
  const int xi = fold_fast((double)x[a], axis(kp, 0));
  const int yi = fold_fast((double)y[a], axis(kp, 1));
  const int zi = fold_fast((double)z[a], axis(kp, 2));
  const uint32_t t = (uint32_t)(xi / g.ts[0] + g.nt[0] * (yi / g.ts[1] + g.nt[1] * (zi / g.ts[2])));
  const uint32_t key = (uint32_t)p * (uint32_t)g.ntiles + t;
  keys[s] = key;
  atomicAdd(&count[key], 1u);
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
  new_slot[s] = start[k] + atomicAdd(&fill[k], 1u);
}

// Move every chain to its new slot (positions double-buffered) and update
// the slot <-> chain maps.  One lane per bead; the slot of a bead follows
// from its architecture block by one fp64 product, corrected by one step.
__global__ void __launch_bounds__(256) k_permute(KParams kp, const float* __restrict__ x,
                                                 const float* __restrict__ y,
                                                 const float* __restrict__ z,
                                                 float* __restrict__ x2, float* __restrict__ y2,
                                                 float* __restrict__ z2,
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

// One workgroup per tile: bin the beads of the tile's chains into the LDS
// window, store the window.  Beads outside the window go straight to the
// final counts with a global atomic.
template <int BITS>
__global__ void __launch_bounds__(1024) k_grid_tiles(const KParams* __restrict__ Ps, int sys,
                                                     TileGeo g,
                                                     const uint32_t* __restrict__ tile_start,
                                                     uint32_t* __restrict__ scratch,
                                                     unsigned long long* __restrict__ fallback) {
  extern __shared__ uint32_t win[];
  KPC& kp = param(Ps, (uint32_t)sys);
  const SCMF_GLOBAL float* __restrict__ x = gptr((const float*)kp.x);
  const SCMF_GLOBAL float* __restrict__ y = gptr((const float*)kp.y);
  const SCMF_GLOBAL float* __restrict__ z = gptr((const float*)kp.z);
  const SCMF_GLOBAL uint8_t* __restrict__ btype = gptr((const uint8_t*)kp.btype);
  SCMF_GLOBAL uint32_t* __restrict__ counts = gptr((uint32_t*)kp.counts);
  uint32_t* __restrict__ overflow = (uint32_t*)kp.overflow;
  const int t = blockIdx.x;
  for (int w = threadIdx.x; w < g.win_words; w += 1024) win[w] = 0u;
  __syncthreads();
  const int tx = t % g.nt[0], ty = (t / g.nt[0]) % g.nt[1], tz = t / (g.nt[0] * g.nt[1]);
  const int ox = tx * g.ts[0] - g.h, oy = ty * g.ts[1] - g.h, oz = tz * g.ts[2] - g.h;
  for (int p = 0; p < kp.P; ++p) {
    const int64_t s0 = tile_start[p * g.ntiles + t], s1 = tile_start[p * g.ntiles + t + 1];
    if (s1 <= s0) continue;
    const int len = kp.len[p];
    const int64_t b0 = kp.bead_base[p] + (s0 - kp.chain_base[p]) * len;
    const int64_t b1 = kp.bead_base[p] + (s1 - kp.chain_base[p]) * len;
    for (int64_t gb = b0 + threadIdx.x; gb < b1; gb += 1024) {
      const uint32_t ty_ = btype[gb];
      const int xi = fold_fast((double)x[gb], axis(kp, 0));
      const int yi = fold_fast((double)y[gb], axis(kp, 1));
      const int zi = fold_fast((double)z[gb], axis(kp, 2));
      const int wx = wrap_rel(xi, ox, kp.nx), wy = wrap_rel(yi, oy, kp.ny),
                wz = wrap_rel(zi, oz, kp.nz);
      bool done = false;
      if (wx < g.W[0] && wy < g.W[1] && wz < g.W[2]) {
        const uint32_t wc = ty_ * (uint32_t)g.win_cells + (uint32_t)(wx + g.W[0] * (wy + g.W[1] * wz));
        const uint32_t sh = (wc & 1u) * 16u;
        const uint32_t old = atomicAdd(&win[wc >> 1], 1u << sh);
        if (((old >> sh) & 0xFFFFu) == 0xFFFFu) {
          atomicSub(&win[wc >> 1], 1u << sh);  // this half is full: undo, count globally
        } else {
          done = true;
        }
      }
      const uint64_t missed = __ballot(!done);
      if (missed && (threadIdx.x & 63) == (uint32_t)(__ffsll((unsigned long long)missed) - 1))
        atomicAdd(fallback, (unsigned long long)__popcll(missed));
      if (!done) {
        const uint32_t cell = ty_ * kp.n_cells + (uint32_t)(xi + kp.nx * (yi + kp.ny * zi));
        if (BITS == 32) {
          atomicAdd((uint32_t*)&counts[cell], 1u);
        } else {
          const uint32_t sh = (cell & 1u) * 16u;
          const uint32_t old = atomicAdd((uint32_t*)&counts[cell >> 1], 1u << sh);
          if (((old >> sh) & 0xFFFFu) == 0xFFFFu) atomicOr(overflow, 1u);
        }
      }
    }
  }
  __syncthreads();
  uint32_t* dst = scratch + (int64_t)t * g.win_words;
  for (int w = threadIdx.x; w < g.win_words; w += 1024) dst[w] = win[w];
}

// Windows of the tiles covering coordinate c along one axis: up to three
// (tile, offset-in-window) pairs.
__device__ __forceinline__ int covering(int c, int ts, int nt, int h, int W, int n, int* tile,
                                        int* off) {
  const int home = c / ts;
  int k = 0;
  for (int d = -1; d <= 1; ++d) {
    int tt = home + d;
    if (tt < 0) tt += nt;
    if (tt >= nt) tt -= nt;
    const int o = wrap_rel(c, tt * ts - h, n);
    if (o < W) {
      bool dup = false;
      for (int j = 0; j < k; ++j) dup |= tile[j] == tt;
      if (!dup) {
        tile[k] = tt;
        off[k] = o;
        ++k;
      }
    }
  }
  return k;
}

__device__ __forceinline__ uint32_t window_count(const uint32_t* scratch, const TileGeo& g, int t,
                                                 uint32_t type, int wx, int wy, int wz) {
  const uint32_t wc = type * (uint32_t)g.win_cells + (uint32_t)(wx + g.W[0] * (wy + g.W[1] * wz));
  return (scratch[(int64_t)t * g.win_words + (wc >> 1)] >> ((wc & 1u) * 16u)) & 0xFFFFu;
}

// Sum the windows into the final counts (which already hold the fallback
// atomics).  One lane per 32-bit word of counts: one cell (uint32) or a pair
// of cells (uint16), so no two lanes write the same word.
template <int BITS>
__global__ void __launch_bounds__(256) k_tiles_sum(const KParams* __restrict__ Ps, int sys,
                                                   TileGeo g,
                                                   const uint32_t* __restrict__ scratch) {
  KPC& kp = param(Ps, (uint32_t)sys);
  SCMF_GLOBAL uint32_t* __restrict__ counts = gptr((uint32_t*)kp.counts);
  uint32_t* __restrict__ overflow = (uint32_t*)kp.overflow;
  const int64_t n_words = kp.count_words;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x; w < n_words; w += stride) {
    uint32_t out = 0;
    const int per_word = BITS == 32 ? 1 : 2;
    for (int h2 = 0; h2 < per_word; ++h2) {
      const int64_t idx = w * per_word + h2;
      if (idx >= (int64_t)kp.T * kp.n_cells) break;
      const uint32_t type = (uint32_t)(idx / kp.n_cells);
      const uint32_t c = (uint32_t)(idx - (int64_t)type * kp.n_cells);
      const uint32_t row = c / (uint32_t)kp.nx;
      const int xi = (int)(c - row * (uint32_t)kp.nx);
      const int yi = (int)(row % (uint32_t)kp.ny), zi = (int)(row / (uint32_t)kp.ny);
      int txs[3], oxs[3], tys[3], oys[3], tzs[3], ozs[3];
      const int kx = covering(xi, g.ts[0], g.nt[0], g.h, g.W[0], kp.nx, txs, oxs);
      const int ky = covering(yi, g.ts[1], g.nt[1], g.h, g.W[1], kp.ny, tys, oys);
      const int kz = covering(zi, g.ts[2], g.nt[2], g.h, g.W[2], kp.nz, tzs, ozs);
      uint32_t sum = BITS == 32 ? counts[w] : (counts[w] >> (16 * h2)) & 0xFFFFu;
      for (int a = 0; a < kz; ++a)
        for (int b = 0; b < ky; ++b)
          for (int d = 0; d < kx; ++d) {
            const int t = txs[d] + g.nt[0] * (tys[b] + g.nt[1] * tzs[a]);
            sum += window_count(scratch, g, t, type, oxs[d], oys[b], ozs[a]);
          }
      if (BITS == 32) {
        out = sum;
      } else {
        if (sum > 0xFFFFu) {
          atomicOr(overflow, 1u);
          sum = 0xFFFFu;
        }
        out |= sum << (16 * h2);
      }
    }
    counts[w] = out;
  }
}

}  // namespace scmf
