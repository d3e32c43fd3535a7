// scmf_engine.hip — HIP kernels for gfx950 and the C ABI of include/scmf.h.
//
// One system lives on one device.  Device layout:
//   x[], y[], z[]  fp32, n_beads each; chain c owns beads
//                  [bead_base[p] + (c - chain_base[p]) N_p, ... + N_p)
//   counts         uint16 pairs packed in uint32 words, or uint32, [type][z][y][x]
//   omega          fp32 [type][z][y][x]
// Kernels:
//   k_init_coords  one lane per chain: start point + Gaussian walk
//   k_sweep_lds    one lane per chain, the wavefront's 64 chains staged in LDS:
//                  N sequential smart-MC / trial moves per chain
//   k_sweep_global the same with positions in global memory (chains too long
//                  for an LDS tile)
//   k_zero         clear the counts (a kernel, not a memset, so it can sit in a graph)
//   k_grid         one lane per bead: NGP cell, atomic increment
//   k_omega        one lane per cell: all T omega values in fp64, stored fp32
//   k_omega4       the same, four cells per lane with vector loads and stores
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scmf.h"
#include "scmf_batch_tickets.h"
#include "scmf_device.h"
#include "scmf_host.h"
#include "scmf_tiles.h"

#pragma clang fp contract(off)

using namespace scmf;

// ============================================================== kernels
namespace {

constexpr int BLOCK = 256;
constexpr int MAX_INIT_N = 4096;  // occupancy bitmask of the initial walk

// initialize_coord, device_function.cu:140-207, one lane per chain.
__global__ void __launch_bounds__(BLOCK) k_init_coords(const KParams* __restrict__ Ps,
                                                       const BlockRef* __restrict__ bt) {
  const BlockRef br = block_ref(bt);
  KPC& kp = param(Ps, br.sys);
  SCMF_GLOBAL float* __restrict__ x = gptr((float*)kp.x);
  SCMF_GLOBAL float* __restrict__ y = gptr((float*)kp.y);
  SCMF_GLOBAL float* __restrict__ z = gptr((float*)kp.z);
  const int64_t chain = (int64_t)br.local * BLOCK + threadIdx.x;
  if (chain >= kp.n_chains) return;
  const int p = arch_of_chain(kp, chain);
  const int N = kp.len[p];
  const int64_t pi = kp.bead_base[p] + (chain - kp.chain_base[p]) * N;
  uint32_t occ[MAX_INIT_N / 32];
  for (int w = 0; w < (N + 31) / 32; ++w) occ[w] = 0;
  const u32x4 w0 = stream_block(kp.seed, kp.system_id, (uint32_t)chain, INIT_SWEEP, 0, SLOT_INIT);
  x[pi] = (float)((double)u01_24f(w0.v[0]) * kp.Lx);
  y[pi] = (float)((double)u01_24f(w0.v[1]) * kp.Ly);
  z[pi] = (float)((double)u01_24f(w0.v[2]) * kp.Lz);
  occ[0] = 1u;
  const double cste = 1.0 / sqrt(3.0 * ((double)N - 1.0));
  const int mo = kp.mono_off[p];
  for (int i = 0; i < N; ++i) {
    for (int k = kp.nbr_start[mo + i]; k < kp.nbr_start[mo + i + 1]; ++k) {
      const int j = kp.nbr[k];
      if (((occ[j >> 5] >> (j & 31)) & 1u) == 0) {
        const u32x4 w =
            stream_block(kp.seed, kp.system_id, (uint32_t)chain, INIT_SWEEP, (uint32_t)(1 + j), SLOT_INIT);
        double n0, n1, n2;
        normal3(w, &n0, &n1, &n2);
        x[pi + j] = (float)(n0 * cste + (double)x[pi + i]);
        y[pi + j] = (float)(n1 * cste + (double)y[pi + i]);
        z[pi + j] = (float)(n2 * cste + (double)z[pi + i]);
        occ[j >> 5] |= 1u << (j & 31);
      }
    }
  }
}

// Positions of one chain, seen by one lane.
struct GlobalChain {  // planar global arrays, chain starting at `base`
  SCMF_GLOBAL float* x;
  SCMF_GLOBAL float* y;
  SCMF_GLOBAL float* z;
  __device__ float X(int b) const { return x[b]; }
  __device__ float Y(int b) const { return y[b]; }
  __device__ float Z(int b) const { return z[b]; }
  __device__ void set(int b, float a, float c, float d) const {
    x[b] = a;
    y[b] = c;
    z[b] = d;
  }
};
template <int STRIDE>
struct LdsChain {  // LDS columns: bead b of this lane's chain at [b * STRIDE]
  float* x;
  float* y;
  float* z;
  __device__ float X(int b) const { return x[b * STRIDE]; }
  __device__ float Y(int b) const { return y[b * STRIDE]; }
  __device__ float Z(int b) const { return z[b * STRIDE]; }
  __device__ void set(int b, float a, float c, float d) const {
    x[b * STRIDE] = a;
    y[b * STRIDE] = c;
    z[b * STRIDE] = d;
  }
};

// Calls f(j) for every neighbour j of bead b, ascending (the order of
// init_config.cu:330-336): computed for linear chains, looked up otherwise.
template <class KP, class F>
__device__ __forceinline__ void for_neighbours(const KP& kp, int p, int b, int len, F&& f) {
  if (kp.linear[p]) {
    if (b > 0) f(b - 1);
    if (b < len - 1) f(b + 1);
  } else {
    const SCMF_GLOBAL int* ns = gptr((const int*)kp.nbr_start);
    const SCMF_GLOBAL int* nb = gptr((const int*)kp.nbr);
    const int k0 = ns[kp.mono_off[p] + b], k1 = ns[kp.mono_off[p] + b + 1];
    for (int k = k0; k < k1; ++k) f(nb[k]);
  }
}

constexpr int MAX_DEG = 8;  // bonds per bead handled by the sweep kernels

// Bonded neighbours of a bead, in ascending bead index.  For linear chains
// (bonds i <-> i+1) the two possible neighbours are read once per move into
// registers; otherwise they are read where used.
template <bool LINEAR, class KP, class V>
struct Neighbours;

template <class KP, class V>
struct Neighbours<true, KP, V> {
  bool hl, hr;  // bead b-1, bead b+1 exist
  float lx, ly, lz, rx, ry, rz;
  __device__ Neighbours(const KP&, const V& c, int, int b, int len) : hl(b > 0), hr(b < len - 1) {
    lx = hl ? c.X(b - 1) : 0.0f;
    ly = hl ? c.Y(b - 1) : 0.0f;
    lz = hl ? c.Z(b - 1) : 0.0f;
    rx = hr ? c.X(b + 1) : 0.0f;
    ry = hr ? c.Y(b + 1) : 0.0f;
    rz = hr ? c.Z(b + 1) : 0.0f;
  }
  template <class F>
  __device__ void each(F&& f) const {
    if (hl) f(lx, ly, lz);
    if (hr) f(rx, ry, rz);
  }
};

template <class KP, class V>
struct Neighbours<false, KP, V> {
  const KP& kp;
  const V& c;
  int moff, b;
  __device__ Neighbours(const KP& kp_, const V& c_, int moff_, int b_, int)
      : kp(kp_), c(c_), moff(moff_), b(b_) {}
  template <class F>
  __device__ void each(F&& f) const {
    const SCMF_GLOBAL int* ns = gptr((const int*)kp.nbr_start);
    const SCMF_GLOBAL int* nb = gptr((const int*)kp.nbr);
    const int k0 = ns[moff + b], k1 = ns[moff + b + 1];
    for (int k = k0; k < k1; ++k) {
      const int j = nb[k];
      f(c.X(j), c.Y(j), c.Z(j));
    }
  }
};

// What the moves of one chain read about its architecture, loaded once per
// chain before its moves (the loads would otherwise open every move, one
// waiting on the next): the chain length, whether its bonds are i <-> i+1,
// where its types and neighbour lists start, and its types as a bit mask
// when the system has at most two types and 64 beads a chain.
struct ChainArch {
  int length, moff;
  bool linear;
  uint64_t tbits;
};

template <class KP>
__device__ __forceinline__ ChainArch chain_arch(const KP& kp, int p) {
  ChainArch a;
  a.length = kp.len[p];
  a.moff = kp.mono_off[p];
  a.linear = kp.linear[p] != 0;
  a.tbits = kp.tbits[p];
  return a;
}

// The monomer type of bead b (monotype, init_config.cu:338).
template <class KP>
__device__ __forceinline__ uint32_t bead_type(const KP& kp, const ChainArch& a, int b) {
  if (kp.type_bits) return (uint32_t)(a.tbits >> b) & 1u;
  return (uint32_t)gptr((const int*)kp.monotype)[a.moff + b];
}

// A[t] and R[t] chosen among the system's values (uniform, in scalar
// registers) instead of loaded by a per-lane index.
template <class KP>
__device__ __forceinline__ void smart_step(const KP& kp, uint32_t t, float* A, float* R) {
  float a = kp.A[0], r = kp.R[0];
#pragma unroll
  for (int k = 1; k < MAX_T; ++k) {
    if (k >= kp.T) break;
    if (t == (uint32_t)k) {
      a = kp.A[k];
      r = kp.R[k];
    }
  }
  *A = a;
  *R = r;
}

// add_bond_forces_arr, device_function.cu:589-634: neighbour difference in
// fp32, scaled and accumulated in fp64, added to the fp32 force.  (d * 2.0)
// * harm rounds exactly as d * (2 harm), doubling being exact.
template <class N>
__device__ __forceinline__ void bond_force(const N& nb, double harm2, float px, float py, float pz,
                                           float* fx, float* fy, float* fz) {
  double v1x = 0.0, v1y = 0.0, v1z = 0.0;
  nb.each([&](float qx, float qy, float qz) {
    v1x += (double)(qx - px) * harm2;
    v1y += (double)(qy - py) * harm2;
    v1z += (double)(qz - pz) * harm2;
  });
  *fx = (float)((double)*fx + v1x);
  *fy = (float)((double)*fy + v1y);
  *fz = (float)((double)*fz + v1z);
}

// One move of mc_polymer_move_arr (device_function.cu:747-790) with
// trial_move_smc_arr (:637-717), calc_delta_energy_arr (:562-588),
// calc_delta_bonded_energy_arr (:509-561) and som_accept (:444-454) inlined.
// Move i of the sweep on bead ibead (its pick, drawn by the caller from the
// move's SLOT_PICK block).  Returns 1 if accepted.
template <bool LINEAR, class KP, class V>
__device__ __forceinline__ uint32_t one_move_n(const KP& kp, const V& c, const ChainArch& a,
                                               uint32_t chain, uint32_t sweep, uint32_t i,
                                               int ibead,
                                               const SCMF_GLOBAL float* __restrict__ omega) {
  const uint32_t iwtype = bead_type(kp, a, ibead);
  const u32x4 disp = stream_block(kp.seed, kp.system_id, chain, sweep, i, SLOT_DISP);
  const float ox = c.X(ibead), oy = c.Y(ibead), oz = c.Z(ibead);
  const Neighbours<LINEAR, KP, V> nb(kp, c, LINEAR ? 0 : a.moff, ibead, a.length);
  const double harm2 = kp.harm * 2.0;
  float dx, dy, dz;
  double smc_deltaE = 0.0;
  if (kp.move_type == SCMF_MOVE_SMART) {
    float px = ox, py = oy, pz = oz;
    float A, R;
    smart_step(kp, iwtype, &A, &R);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    bond_force(nb, harm2, px, py, pz, &fx, &fy, &fz);
    float rx, ry, rz;
    normal3f(disp, &rx, &ry, &rz);
    dx = A * fx + rx * R;
    dy = A * fy + ry * R;
    dz = A * fz + rz * R;
    px += dx;
    py += dy;
    pz += dz;
    float nfx = 0.0f, nfy = 0.0f, nfz = 0.0f;
    bond_force(nb, harm2, px, py, pz, &nfx, &nfy, &nfz);
    smc_deltaE = 0.0;
    smc_deltaE += 0.5 * ((nfx + fx) * dx + (nfy + fy) * dy + (nfz + fz) * dz);
    smc_deltaE +=
        0.25 * A * ((nfx * nfx) + (nfy * nfy) + (nfz * nfz) - (fx * fx) - (fy * fy) - (fz * fz));
  } else {
    const float delta = kp.trial_delta;
    dx = delta * (2.0f * u01_24f_closed0(disp.v[0]) - 1.0f);
    dy = delta * (2.0f * u01_24f_closed0(disp.v[1]) - 1.0f);
    dz = delta * (2.0f * u01_24f_closed0(disp.v[2]) - 1.0f);
  }
  const float newx = ox + dx, newy = oy + dy, newz = oz + dz;
  const double xold = ox, yold = oy, zold = oz;
  // old cell from the stored fp32 position, new cell from the fp64 sum (:573-577)
  const uint32_t cell_old = cell_of(kp, ox, oy, oz);
  const uint32_t cell_new = cell_of(kp, xold + (double)dx, yold + (double)dy, zold + (double)dz);
  const double energy_old = omega[iwtype * kp.n_cells + cell_old];
  const double energy_new = omega[iwtype * kp.n_cells + cell_new];
  double bonded = 0;
  {
    // ox - qx is exactly -(qx - ox), the difference the old force used (the
    // sign of a zero difference cannot reach the squares)
    const double ddx = dx, ddy = dy, ddz = dz;
    nb.each([&](float qx, float qy, float qz) {
      const double old_rx = -(double)(qx - ox);
      const double new_rx = old_rx + ddx;
      const double old_ry = -(double)(qy - oy);
      const double new_ry = old_ry + ddy;
      const double old_rz = -(double)(qz - oz);
      const double new_rz = old_rz + ddz;
      const double old_r2 = old_rx * old_rx + old_ry * old_ry + old_rz * old_rz;
      const double new_r2 = new_rx * new_rx + new_ry * new_ry + new_rz * new_rz;
      bonded += kp.harm * (new_r2 - old_r2);
    });
  }
  double energy = energy_new - energy_old;
  energy += bonded;
  double delta_energy = energy;
  delta_energy += smc_deltaE;
  const double p_acc = exp_cw(-1.0 * delta_energy);
  if ((p_acc > 1) || (p_acc > (double)accept_u01(disp))) {
    c.set(ibead, newx, newy, newz);
    return 1u;
  }
  return 0u;
}

// The move with neighbour storage sized for the chain's architecture.
template <class KP, class V>
__device__ __forceinline__ uint32_t one_move(const KP& kp, const V& c, const ChainArch& a,
                                             uint32_t chain, uint32_t sweep, uint32_t i, int ibead,
                                             const SCMF_GLOBAL float* __restrict__ omega) {
  if (a.linear) return one_move_n<true>(kp, c, a, chain, sweep, i, ibead, omega);
  return one_move_n<false>(kp, c, a, chain, sweep, i, ibead, omega);
}

// The N moves of one chain in order, on one lane: one pick block per two moves.
template <class KP, class V>
__device__ __forceinline__ uint32_t chain_moves(const KP& kp, const V& c, int p, int length,
                                                uint32_t chain, uint32_t sweep,
                                                const SCMF_GLOBAL float* __restrict__ omega) {
  const ChainArch a = chain_arch(kp, p);
  uint32_t acc = 0;
  u32x4 pk{};
  for (int i = 0; i < length; i++) {
    if ((i & 1) == 0) pk = stream_block(kp.seed, kp.system_id, chain, sweep, (uint32_t)i >> 1, SLOT_PICK);
    const int ib = pick_bead(pick_u01(pk, (uint32_t)i), length);
    acc += one_move(kp, c, a, chain, sweep, (uint32_t)i, ib, omega);
  }
  return acc;
}

__device__ __forceinline__ void wave_count(uint32_t acc, unsigned long long* accepts) {
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0 && acc)
    atomicAdd(&accepts[(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (ACC_SLOTS - 1)],
              (unsigned long long)acc);
}

// Advance the stream sweep word of systems [0, n) after their sweep; a
// member at its stop stays, and ctr[1] records whether the sweep ran.
__global__ void __launch_bounds__(64) k_advance(const KParams* __restrict__ Ps, int n) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i < n) {
    uint32_t* c = Ps[i].ctr;
    const bool ran = c[0] < Ps[i].stop;
    if (ran) c[0] = c[0] + 1u;
    c[1] = ran ? 1u : 0u;
  }
}

// mc_polymer_move_arr, one lane per chain, positions read and written in
// global memory (used when a chain tile does not fit in LDS).
__global__ void __launch_bounds__(BLOCK) k_sweep_global(const KParams* __restrict__ Ps,
                                                        const BlockRef* __restrict__ bt, int clear,
                                                        int64_t c0, int64_t c1) {
  const BlockRef br = block_ref(bt);
  KPC& kp = param(Ps, br.sys);
  if (sweep_stopped(kp)) return;
  if (clear) clear_counts(kp);
  SCMF_GLOBAL float* __restrict__ x = gptr((float*)kp.x);
  SCMF_GLOBAL float* __restrict__ y = gptr((float*)kp.y);
  SCMF_GLOBAL float* __restrict__ z = gptr((float*)kp.z);
  const SCMF_GLOBAL float* __restrict__ omega = gptr((const float*)kp.omega);
  const uint32_t sweep = ((const uint32_t*)kp.ctr)[0];
  const int64_t chain = c0 + (int64_t)br.local * BLOCK + threadIdx.x;
  uint32_t acc = 0;
  if (chain < c1 && chain < kp.n_chains) {
    const int p = arch_of_chain(kp, chain);
    const int length = kp.len[p];
    const int64_t off = kp.bead_base[p] + (chain - kp.chain_base[p]) * length;
    const GlobalChain c{x + off, y + off, z + off};
    const uint32_t id = gptr((const uint32_t*)kp.slot_chain)[chain];
    acc = chain_moves(kp, c, p, length, id, sweep, omega);
  }
  wave_count(acc, (unsigned long long*)kp.accepts);
}

// mc_polymer_move_arr with the wavefront's 64 chains staged in LDS: one lane
// per chain, bead b of lane l at LDS word b*64 + l (conflict-free for any
// per-lane bead index).  Each lane touches only its own column, so no barrier
// is needed; global memory sees one read and one write of every coordinate.
__global__ void __launch_bounds__(64) k_sweep_lds(const KParams* __restrict__ Ps,
                                                  const BlockRef* __restrict__ bt, int nmax,
                                                  int clear, int64_t c0, int64_t c1) {
  extern __shared__ float lds[];
  const BlockRef br = block_ref(bt);
  KPC& kp = param(Ps, br.sys);
  if (sweep_stopped(kp)) return;
  if (clear) clear_counts(kp);
  SCMF_GLOBAL float* __restrict__ x = gptr((float*)kp.x);
  SCMF_GLOBAL float* __restrict__ y = gptr((float*)kp.y);
  SCMF_GLOBAL float* __restrict__ z = gptr((float*)kp.z);
  const SCMF_GLOBAL float* __restrict__ omega = gptr((const float*)kp.omega);
  const uint32_t sweep = ((const uint32_t*)kp.ctr)[0];
  const int lane = threadIdx.x;
  const int64_t chain = c0 + (int64_t)br.local * 64 + lane;
  uint32_t acc = 0;
  if (chain < c1 && chain < kp.n_chains) {
    const int p = arch_of_chain(kp, chain);
    const int length = kp.len[p];
    const int64_t off = kp.bead_base[p] + (chain - kp.chain_base[p]) * length;
    float* sx = lds + lane;
    float* sy = lds + nmax * 64 + lane;
    float* sz = lds + 2 * nmax * 64 + lane;
    for (int b = 0; b < length; ++b) {
      sx[b * 64] = x[off + b];
      sy[b * 64] = y[off + b];
      sz[b * 64] = z[off + b];
    }
    const LdsChain<64> c{sx, sy, sz};
    const uint32_t id = gptr((const uint32_t*)kp.slot_chain)[chain];
    acc = chain_moves(kp, c, p, length, id, sweep, omega);
    for (int b = 0; b < length; ++b) {
      x[off + b] = sx[b * 64];
      y[off + b] = sy[b * 64];
      z[off + b] = sz[b * 64];
    }
  }
  wave_count(acc, (unsigned long long*)kp.accepts);
}

// mc_polymer_move_arr with L lanes per chain (G = 64/L chains per wavefront).
//
// The N moves of a chain in one sweep are not all dependent: move i reads the
// closed neighbourhood {b_i} + bonded(b_i) and writes b_i, so it depends on an
// earlier move j only if b_j lies in that neighbourhood.  The picks b_i come
// from the counter streams before any move runs, so each chain's moves are
// put into rounds first: move i goes to the first round after the rounds of
// the earlier moves that wrote a bead of its neighbourhood (the last writer of
// each bead has the latest such round) that holds fewer than L moves.  Moves
// of one round are independent and run on the chain's L lanes side by side,
// rounds in order.  Any order that keeps every pair of dependent moves in
// sequence gives every move exactly the state it sees in the sequential
// order, so results are bit-identical to one lane per chain; the per-wavefront
// LDS tile shrinks by L.  (Packing rounds to L moves instead of running
// dependency levels one after the other, ceil(size / L) passes each, cuts
// the passes of a wave of 8 chains of 64 beads from ~13.1 to ~10.1: lane use
// 61% -> 80%, by simulation of the picks.)
//
// LDS per chain (index type I: uint8 for N <= 255): positions 3N floats
// (bead b at [b*G + slot]); lev[N] (the bead each move picks); W[N+2]
// (round per bead during the scan, bead b at W[b+1] with zero sentinels
// around, then per-round fill counts); start[N+2] (moves per round during
// the scan, then the first list entry of each round); list[N] (moves by
// round); rnd[N] (the round of each move).  With a schedule made by
// the schedule kernels the book keeps start, list and the bead of each move.
// (Five waves per SIMD, at most 102 VGPRs: the LDS tiles allow about as
// many, and the move is issue-bound, so occupancy is worth registers.)
// Byte k of the move-schedule record of a slot (records interleaved by
// groups of 8 slots; made by k_schedule_reg or k_schedule_lds below).
__device__ __forceinline__ int64_t sched_index(int64_t slot, int stride, int k) {
  return ((slot >> 3) * (int64_t)stride + k) * 8 + (slot & 7);
}

template <int L, typename I>
__global__ void __launch_bounds__(64, 5) k_sweep_sched(const KParams* __restrict__ Ps,
                                                    const BlockRef* __restrict__ bt, int nmax,
                                                    int clear, int64_t c0, int64_t c1) {
  constexpr int G = 64 / L;
  extern __shared__ float lds[];
  const BlockRef br = block_ref(bt);
  KPC& kp = param(Ps, br.sys);
  if (sweep_stopped(kp)) return;
  if (clear) clear_counts(kp);
  SCMF_GLOBAL float* __restrict__ x = gptr((float*)kp.x);
  SCMF_GLOBAL float* __restrict__ y = gptr((float*)kp.y);
  SCMF_GLOBAL float* __restrict__ z = gptr((float*)kp.z);
  const SCMF_GLOBAL float* __restrict__ omega = gptr((const float*)kp.omega);
  const uint32_t sweep = ((const uint32_t*)kp.ctr)[0];
  const int lane = threadIdx.x;
  const int slot = lane / L, r = lane % L;
  float* sx = lds;
  float* sy = lds + nmax * G;
  float* sz = lds + 2 * nmax * G;
  I* book = (I*)(lds + 3 * nmax * G);
  // A schedule made by the schedule kernels (single systems) needs start and list
  // only; otherwise the chain's lanes make it here from lev and W.
  const SCMF_GLOBAL uint8_t* __restrict__ sched = gptr((const uint8_t*)kp.sched);
  const bool pre = sched != nullptr;
  const int per = pre ? 3 * nmax + 4 : 5 * nmax + 4;
  I* lev = book + slot * per;
  I* W = lev + nmax;
  I* start = pre ? lev : W + nmax + 2;
  I* list = start + nmax + 2;
  I* rnd = list + nmax;   // the round of each move (in-sweep scan only)
  I* bead = list + nmax;  // the bead of each move (schedules from the schedule kernels)
  const int64_t chain = c0 + (int64_t)br.local * G + slot;
  const bool active = chain < c1 && chain < kp.n_chains;
  int p = 0, length = 0;
  int64_t off = 0;
  uint32_t id = 0;
  if (active) {
    p = arch_of_chain(kp, chain);
    length = kp.len[p];
    off = kp.bead_base[p] + (chain - kp.chain_base[p]) * length;
    id = gptr((const uint32_t*)kp.slot_chain)[chain];
  }
  // the wave's eight chains are one record group (records interleave groups
  // of 8 slots), all nmax beads long: their entries load 8 bytes at a time
  bool group = false;
  if constexpr (L == 8)
    group = ((c0 + (int64_t)br.local * G) & 7) == 0 && __all(active && length == nmax);
  // For records of at most 256 entries the group's loads are issued before
  // the tile's, so both wait together (four 8-byte entries a lane).
  const int st = kp.sched_stride;
  const bool early = pre && group && 3 * nmax + 2 <= 256;
  uint64_t ge[4] = {};
  int maxl = 0;
  const SCMF_GLOBAL uint64_t* grp = nullptr;
  if (pre && group)
    grp = (const SCMF_GLOBAL uint64_t*)(sched + ((c0 + (int64_t)br.local * G) >> 3) * (int64_t)st * 8);
  if (early) {
    if (active) maxl = sched[sched_index(chain, st, 0)];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 1 + lane + 64 * j;
      if (k <= 3 * nmax + 2) ge[j] = grp[k];
    }
  }
  // load the tile, eight beads a lane in flight at once (the loop with one
  // bead an iteration waited for each: 0.18 ms more a sweep at 268M beads)
  for (int b0 = r; b0 < length; b0 += 8 * L) {
    float tx[8], ty[8], tz[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int b = b0 + u * L;
      if (b < length) {
        tx[u] = x[off + b];
        ty[u] = y[off + b];
        tz[u] = z[off + b];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int b = b0 + u * L;
      if (b < length) {
        sx[b * G + slot] = tx[u];
        sy[b * G + slot] = ty[u];
        sz[b * G + slot] = tz[u];
      }
    }
  }
  if (pre) {
    if (early) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 1 + lane + 64 * j;
        if (k <= 3 * nmax + 2) {
#pragma unroll
          for (int q = 0; q < 8; ++q) book[q * per + k - 1] = (I)((ge[j] >> (8 * q)) & 0xFFu);
        }
      }
    } else if (group) {
      // the group's entries 1 .. 3N+2 (start, list, bead: book offsets 0 ..
      // 3N+1 of each slot), one 8-byte load per entry for all eight slots
      // (0.15 ms less a sweep at 268M beads than a byte load per lane and
      // entry; the same treatment of the tile's loads gained nothing, and of
      // its write-back and cells cost 0.35 ms)
      if (active) maxl = sched[sched_index(chain, st, 0)];
      for (int k = 1 + threadIdx.x; k <= 3 * nmax + 2; k += 64) {
        const uint64_t e = grp[k];
#pragma unroll
        for (int j = 0; j < 8; ++j) book[j * per + k - 1] = (I)((e >> (8 * j)) & 0xFFu);
      }
    } else if (active) {
      maxl = sched[sched_index(chain, st, 0)];
      for (int k = r; k <= length + 1; k += L) start[k] = (I)sched[sched_index(chain, st, 1 + k)];
      for (int k = r; k < length; k += L) {
        list[k] = (I)sched[sched_index(chain, st, length + 3 + k)];
        bead[k] = (I)sched[sched_index(chain, st, 2 * length + 3 + k)];
      }
    }
    __syncthreads();
  } else {
  for (int i = 2 * r; i < length; i += 2 * L) {  // one pick block per two moves
    const u32x4 pick = stream_block(kp.seed, kp.system_id, id, sweep, (uint32_t)i >> 1, SLOT_PICK);
    lev[i] = (I)pick_bead(pick_u01(pick, (uint32_t)i), length);
    if (i + 1 < length) lev[i + 1] = (I)pick_bead(pick_u01(pick, (uint32_t)i + 1), length);
  }
  for (int k = r; k <= length + 1; k += L) W[k] = 0;
  for (int k = r; k <= length + 1; k += L) start[k] = 0;
  __syncthreads();
  // rounds (one lane per chain), round sizes, moves listed by round
  if (r == 0 && active) {
    I* const Wb = W + 1;  // level of each bead, with zero sentinels at beads -1 and N
    int maxl = 0;
    if (kp.linear[p]) {
      // bonds i <-> i+1: the neighbourhood of b is b-1, b, b+1 (sentinels at the ends)
      for (int i = 0; i < length; ++i) {
        const int b = lev[i];
        int l = max(max((int)Wb[b - 1], (int)Wb[b]), (int)Wb[b + 1]) + 1;
        while ((int)start[l + 1] >= L) ++l;  // round full: the next one
        Wb[b] = (I)l;
        rnd[i] = (I)l;
        start[l + 1] = (I)(start[l + 1] + 1);
        maxl = max(maxl, l);
      }
    } else {
      for (int i = 0; i < length; ++i) {
        const int b = lev[i];
        int l = Wb[b];
        for_neighbours(kp, p, b, length, [&](int j) { l = max(l, (int)Wb[j]); });
        l += 1;
        while ((int)start[l + 1] >= L) ++l;
        Wb[b] = (I)l;
        rnd[i] = (I)l;
        start[l + 1] = (I)(start[l + 1] + 1);
        maxl = max(maxl, l);
      }
    }
    for (int l = 1; l <= maxl; ++l) {
      start[l + 1] = (I)(start[l + 1] + start[l]);
      W[l] = 0;
    }
    for (int i = 0; i < length; ++i) {
      const int l = rnd[i];
      const int k = W[l];
      W[l] = (I)(k + 1);
      list[start[l] + k] = (I)i;
    }
    W[0] = (I)maxl;
  }
  __syncthreads();
  maxl = active ? (int)W[0] : 0;
  }
  uint32_t acc = 0;
  const LdsChain<G> c{sx + slot, sy + slot, sz + slot};
  const ChainArch ca = chain_arch(kp, p);
  int l = 1, j = 0;
  while (__any(l <= maxl)) {
    if (l <= maxl) {
      const int s0 = start[l], cl = (int)start[l + 1] - s0;
      const int k = j * L + r;
      if (k < cl) {
        const int i = list[s0 + k];
        // the bead the schedule kernel picked for this move (no pick block
        // drawn here)
        const int ib = pre ? (int)bead[i] : (int)lev[i];
        acc += one_move(kp, c, ca, id, sweep, (uint32_t)i, ib, omega);
      }
      if (++j * L >= cl) {
        ++l;
        j = 0;
      }
    }
    __syncthreads();
  }
  // write the chain back, with the packed cell of every bead for the gridding
  SCMF_GLOBAL uint32_t* __restrict__ cells = gptr(kp.cells);
  if (cells) {
    const Axis ax0 = axis(kp, 0), ax1 = axis(kp, 1), ax2 = axis(kp, 2);
    for (int b = r; b < length; b += L) {
      const float px = sx[b * G + slot], py = sy[b * G + slot], pz = sz[b * G + slot];
      x[off + b] = px;
      y[off + b] = py;
      z[off + b] = pz;
      int xi, yi, zi;
      const bool quick = kp.pow2 && (fold_pow2(px, ax0, &xi) & fold_pow2(py, ax1, &yi) &
                                     fold_pow2(pz, ax2, &zi));
      if (!quick) fold3_general(px, py, pz, ax0, ax1, ax2, &xi, &yi, &zi);
      cells[off + b] = pack_cell(xi, yi, zi);
    }
  } else {
    for (int b = r; b < length; b += L) {
      x[off + b] = sx[b * G + slot];
      y[off + b] = sy[b * G + slot];
      z[off + b] = sz[b * G + slot];
    }
  }
  wave_count(acc, (unsigned long long*)kp.accepts);
}

template <int L, typename I>
size_t sched_lds_bytes(int nmax) {
  constexpr int G = 64 / L;
  return (size_t)3 * nmax * G * sizeof(float) + (size_t)G * (5 * nmax + 4) * sizeof(I);
}

// The same kernel reading a schedule made by the schedule kernels.
size_t sched_pre_lds_bytes(int L, int nmax) {
  const int G = 64 / L;
  return (size_t)3 * nmax * G * sizeof(float) + (size_t)G * (3 * nmax + 4);
}

// The move schedule of every chain for the coming sweep, one lane per slot
// (k_sweep_sched makes the same schedule with one lane of each chain's L,
// the others idle), in rounds of at most `cap` moves (the sweep's lanes per
// chain).  The picks depend on the streams only, so the schedule does not
// wait for positions.  Record of a slot (uint8, chains of at most 255
// beads): entry 0 the number of rounds, entries 1 .. N+2 start[0 .. N+1]
// (the first list entry of each round, through start[maxl + 1]; the entries
// after it are not read), entries
// N+3 .. 2N+2 the move indices by round, entries 2N+3 .. 3N+2 the bead each
// move picked (by move index; the sweep then draws no pick block).  Records are interleaved by groups
// of 8 slots (the chains of one sweep wavefront at 8 lanes per chain): entry
// k of slot s is byte ((s / 8) * stride + k) * 8 + s % 8, so a sweep
// wavefront reads its chains' records with contiguous loads.  The lane
// arrays of these kernels live in LDS interleaved by lane (entry k of lane t
// at k * 64 + t).  The scan is bound by the latency of its LDS round trips,
// so LDS per chain (waves per CU) decides its speed.
constexpr int SCHED_LANES = 64;
size_t schedule_record_bytes(int nmax, int64_t chains) {
  return (size_t)(3 * nmax + 3) * SCHED_LANES * (size_t)((chains + SCHED_LANES - 1) / SCHED_LANES);
}
// Whether the register-count kernel takes chains of nmax beads in rounds of cap moves.
bool schedule_in_registers(int nmax, int cap) {
  return nmax <= 128 && (cap == 2 || cap == 4 || cap == 8 || cap == 16);
}
size_t schedule_lds_bytes(int nmax, int cap) {
  return (size_t)(schedule_in_registers(nmax, cap) ? 2 * nmax + 2 : 4 * nmax + 5) * SCHED_LANES;
}

// Chains of at most 64 NW beads (NW = 1 or 2), rounds of at most cap =
// 2^CAPLOG moves.  The number of moves in each round is kept in registers as
// CAPLOG + 1 bit-planes of 64 NW bits (bit l - 1 of plane j = bit j of round
// l's count), so a move finds the first round with room by a bit scan of the
// top plane (the full rounds) and counts itself by a ripple over the planes:
// the scan waits for LDS once per move (its neighbours' rounds).  The rounds
// used are always 1 .. max, at most one per move, so 64 NW bits hold them
// (the two words of a plane are chosen by selects, never by a register
// index, which would put the planes in scratch memory).  LDS
// holds the round of each bead and of each move (2N + 2 bytes a chain); the
// record goes to memory byte by byte, the bead of each move as the scan
// picks it (staging the record in LDS for 16-byte stores costs waves, and
// measured slower).
template <int CAPLOG, int NW>
__global__ void __launch_bounds__(SCHED_LANES) k_schedule_reg(const KParams* __restrict__ Ps,
                                                              int nmax, int64_t c0, int64_t c1) {
  extern __shared__ uint8_t sh[];
  KPC& kp = param(Ps, 0);
  const int t = threadIdx.x;
  const int64_t chain = c0 + (int64_t)blockIdx.x * SCHED_LANES + t;
  if (chain >= c1) return;
  const uint32_t sweep = ((const uint32_t*)kp.ctr)[0];
  const int p = arch_of_chain(kp, chain);
  const int length = kp.len[p];
  const uint32_t id = gptr((const uint32_t*)kp.slot_chain)[chain];
  const int stride = kp.sched_stride;
  uint8_t* const W = sh + t;                            // [nmax + 2], W[b + 1] = round of bead b
  uint8_t* const lev = W + (nmax + 2) * SCHED_LANES;    // [nmax] round of each move
  auto at = [](uint8_t* a, int k) -> uint8_t& { return a[k * SCHED_LANES]; };
  SCMF_GLOBAL uint8_t* __restrict__ rec = gptr(kp.sched);
  for (int k = 0; k <= length + 1; ++k) at(W, k) = 0;
  const bool linear = kp.linear[p] != 0;
  int maxl = 0;
  uint64_t plane[CAPLOG + 1] = {}, plane_hi[CAPLOG + 1] = {};  // plane_hi: rounds 65 .. 128
  const double n53 = (double)length * 0x1p-53;
  // the record's regions (entry k + 1 lies 8 bytes after entry k)
  SCMF_GLOBAL uint8_t* const bead_rec = rec + sched_index(chain, stride, 2 * length + 3);
  SCMF_GLOBAL uint8_t* const list_rec = rec + sched_index(chain, stride, length + 3);
  SCMF_GLOBAL uint8_t* const start_rec = rec + sched_index(chain, stride, 2);
  // the round of move i on bead cur
  auto step = [&](int i, int cur) {
    int l;
    if (linear) {
      l = max(max((int)at(W, cur), (int)at(W, cur + 1)), (int)at(W, cur + 2)) + 1;
    } else {
      l = at(W, cur + 1);
      for_neighbours(kp, p, cur, length, [&](int j) { l = max(l, (int)at(W, j + 1)); });
      l += 1;
    }
    uint64_t carry, carry_hi = 0;
    if constexpr (NW == 1) {
      l = __builtin_ctzll(~plane[CAPLOG] & (~0ull << (l - 1))) + 1;
      carry = 1ull << (l - 1);
    } else {
      const int k = l - 1;
      const uint64_t lo = k < 64 ? ~plane[CAPLOG] & (~0ull << k) : 0ull;
      const uint64_t hi = ~plane_hi[CAPLOG] & (k < 64 ? ~0ull : ~0ull << (k - 64));
      l = (lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll(hi)) + 1;
      const uint64_t bit = 1ull << ((l - 1) & 63);
      carry = l <= 64 ? bit : 0ull;
      carry_hi = l <= 64 ? 0ull : bit;
    }
#pragma unroll
    for (int j = 0; j <= CAPLOG; ++j) {
      const uint64_t c = plane[j] & carry;
      plane[j] ^= carry;
      carry = c;
      if constexpr (NW == 2) {
        const uint64_t ch = plane_hi[j] & carry_hi;
        plane_hi[j] ^= carry_hi;
        carry_hi = ch;
      }
    }
    at(W, cur + 1) = (uint8_t)l;
    at(lev, i) = (uint8_t)l;
    bead_rec[8 * i] = (uint8_t)cur;
    maxl = max(maxl, l);
  };
  // rounds: one pick block per two moves, drawn at the top of each pair
  // (a block drawn under a condition carried across iterations kept the
  // key schedule in vector registers)
  for (int i = 0; i < length; i += 2) {
    const u32x4 blk = stream_block(kp.seed, kp.system_id, id, sweep, (uint32_t)i >> 1, SLOT_PICK);
    const int b1 = pick_of(blk, (uint32_t)i + 1, length, n53);
    step(i, pick_of(blk, (uint32_t)i, length, n53));
    if (i + 1 < length) step(i + 1, b1);
  }
  // round starts (start[0] = start[1] = 0, prefix sums through the last
  // round, zero after it); W[l] becomes round l's next list position
  rec[sched_index(chain, stride, 0)] = (uint8_t)maxl;
  rec[sched_index(chain, stride, 1)] = 0;
  rec[sched_index(chain, stride, 2)] = 0;
  // (only through the last round: the sweep reads round starts up to maxl + 1)
  int run = 0;
  for (int l = 1; l <= maxl; ++l) {
    int n = 0;
#pragma unroll
    for (int j = 0; j <= CAPLOG; ++j) {
      const uint64_t w = NW == 1 || l <= 64 ? plane[j] : plane_hi[j];
      n |= (int)((w >> ((l - 1) & 63)) & 1u) << j;
    }
    at(W, l) = (uint8_t)run;
    run += n;
    start_rec[8 * l] = (uint8_t)run;
  }
  for (int i = 0; i < length; ++i) {
    const int l = at(lev, i);
    const int pos = at(W, l);
    at(W, l) = (uint8_t)(pos + 1);
    list_rec[8 * pos] = (uint8_t)i;
  }
}

// Any chain length up to 255 and any cap: the round counts live in LDS, in
// the record's start entries.  The records of a workgroup's 64 chains (8
// groups) are one contiguous run of 64 * stride bytes: they are made in LDS
// in that layout and stored with 16-byte stores (the buffer holds whole
// runs).  During the scan the round of each move waits in the record's list
// entries, which the list later fills, and the round of each bead in a lane
// array, which then takes the moves' rounds; the bead of each move goes to
// the record as the scan picks it: 4N + 5 bytes a chain.
__global__ void __launch_bounds__(SCHED_LANES) k_schedule_lds(const KParams* __restrict__ Ps,
                                                              int nmax, int cap) {
  extern __shared__ __align__(16) uint8_t sh[];
  KPC& kp = param(Ps, 0);
  const int t = threadIdx.x;
  const int64_t chain = (int64_t)blockIdx.x * SCHED_LANES + t;
  const bool active = chain < kp.n_chains;
  const uint32_t sweep = ((const uint32_t*)kp.ctr)[0];
  const int p = active ? arch_of_chain(kp, chain) : 0;
  const int length = active ? kp.len[p] : 0;
  const uint32_t id = active ? gptr((const uint32_t*)kp.slot_chain)[chain] : 0u;
  const int stride = kp.sched_stride;
  // this lane's record (entry k at byte 8 k), then its lane array
  uint8_t* const rl = sh + (t >> 3) * stride * 8 + (t & 7);
  uint8_t* const W = sh + SCHED_LANES * stride + t;  // [nmax + 2], W[b + 1] = round of bead b
  auto at = [](uint8_t* a, int k) -> uint8_t& { return a[k * SCHED_LANES]; };
  auto st = [&](int k) -> uint8_t& { return rl[8 * (1 + k)]; };         // start[k], k <= N + 1
  auto lv = [&](int i) -> uint8_t& { return rl[8 * (length + 3 + i)]; };  // list entry i
  auto pk = [&](int i) -> uint8_t& { return rl[8 * (2 * length + 3 + i)]; };  // bead of move i
  u32x4 blk{};
  auto pick = [&](int i) {  // one pick block per two moves
    if ((i & 1) == 0) blk = stream_block(kp.seed, kp.system_id, id, sweep, (uint32_t)i >> 1, SLOT_PICK);
    return pick_bead(pick_u01(blk, (uint32_t)i), length);
  };
  for (int k = 0; k <= length + 1; ++k) {
    at(W, k) = 0;
    st(k) = 0;
  }
  // rounds, the next move's pick drawn while this move's LDS reads are in flight
  const bool linear = kp.linear[p] != 0;
  int maxl = 0;
  int b = length > 0 ? pick(0) : 0;
  for (int i = 0; i < length; ++i) {
    const int cur = b;
    int l;
    if (linear) {
      const int w0 = at(W, cur), w1 = at(W, cur + 1), w2 = at(W, cur + 2);
      if (i + 1 < length) b = pick(i + 1);
      l = max(max(w0, w1), w2) + 1;
    } else {
      l = at(W, cur + 1);
      for_neighbours(kp, p, cur, length, [&](int j) { l = max(l, (int)at(W, j + 1)); });
      l += 1;
      if (i + 1 < length) b = pick(i + 1);
    }
    while (st(l + 1) >= cap) ++l;
    st(l + 1) = (uint8_t)(st(l + 1) + 1);
    at(W, cur + 1) = (uint8_t)l;
    lv(i) = (uint8_t)l;  // (the round of move i, until the list is made)
    pk(i) = (uint8_t)cur;
    maxl = max(maxl, l);
  }
  for (int l = 1; l <= maxl; ++l) st(l + 1) = (uint8_t)(st(l + 1) + st(l));
  rl[0] = (uint8_t)maxl;
  // the moves by round: the rounds of the moves to W, then start[l] serves
  // as round l's next position (ending at the old start[l + 1]), restored
  // after by shifting up
  for (int i = 0; i < length; ++i) at(W, i) = lv(i);
  for (int i = 0; i < length; ++i) {
    const int l = at(W, i);
    const int pos = st(l);
    st(l) = (uint8_t)(pos + 1);
    lv(pos) = (uint8_t)i;
  }
  for (int l = maxl; l >= 1; --l) st(l + 1) = st(l);
  if (length > 0) st(1) = 0;
  __syncthreads();
  // the workgroup's run of records (inactive lanes' bytes are never read)
  uint4* __restrict__ dst = (uint4*)(gptr(kp.sched) + (int64_t)blockIdx.x * SCHED_LANES * stride);
  const uint4* src = (const uint4*)sh;
  for (int q = t; q < 4 * stride; q += SCHED_LANES) dst[q] = src[q];
}


// Clear each system's counts (a kernel, not a memset, so it can sit in a graph).
__global__ void __launch_bounds__(BLOCK) k_zero_counts(const KParams* __restrict__ Ps,
                                                       const BlockRef* __restrict__ bt) {
  const BlockRef br = block_ref(bt);
  KPC& kp = param(Ps, br.sys);
  if (step_skipped(kp)) return;
  const int64_t stride = (int64_t)br.nblk * BLOCK;
  SCMF_GLOBAL uint32_t* __restrict__ c = gptr((uint32_t*)kp.counts);
  for (int64_t i = (int64_t)br.local * BLOCK + threadIdx.x; i < kp.count_words; i += stride) c[i] = 0u;
}

__global__ void __launch_bounds__(BLOCK) k_zero(uint32_t* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

// coord_to_density, device_function.cu:479-508, one lane per bead.  A wave's
// lanes hold consecutive beads of one or two chains, so neighbouring lanes
// often hit the same count word: each run of consecutive lanes with the same
// word adds once, from its first lane, with the run's total (both uint16
// halves of the word at once).  A uint16 half overflows when its old value
// plus the run's increment passes 0xFFFF.  (Merging every equal word of the
// wave instead, one ballot round per distinct word, measured slower: with
// cells of Re/8 a wave holds tens of distinct words, and the small melt's
// gridding took 0.041 ms against 0.018 ms with one atomic per bead and
// 0.017 ms with runs.)
__global__ void __launch_bounds__(BLOCK) k_grid(const KParams* __restrict__ Ps,
                                                const BlockRef* __restrict__ bt, int advance) {
  const BlockRef br = block_ref(bt);
  KPC& kp = param(Ps, br.sys);
  if (step_skipped(kp)) return;
  if (advance) advance_sweep_word(kp);
  const SCMF_GLOBAL float* __restrict__ x = gptr((const float*)kp.x);
  const SCMF_GLOBAL float* __restrict__ y = gptr((const float*)kp.y);
  const SCMF_GLOBAL float* __restrict__ z = gptr((const float*)kp.z);
  SCMF_GLOBAL uint32_t* __restrict__ words = gptr((uint32_t*)kp.counts);
  const bool b32 = kp.density_bits == 32;
  const int lane = (int)(threadIdx.x & 63u);
  const int64_t stride = (int64_t)br.nblk * BLOCK;
  for (int64_t g = (int64_t)br.local * BLOCK + threadIdx.x; g < kp.n_beads; g += stride) {
    const uint32_t t = gptr((const uint8_t*)kp.btype)[g];
    const uint32_t cell = bin_cell(kp, x[g], y[g], z[g], t);
    const uint32_t word = b32 ? cell : cell >> 1;
    const bool high = !b32 && (cell & 1u);
    // runs of consecutive lanes with the same word add once, from the run's
    // first lane, with the run's total
    const uint32_t prev = __shfl_up(word, 1, 64);
    const unsigned long long act = __ballot(1);
    const bool head = lane == 0 || !((act >> (lane - 1)) & 1ull) || prev != word;
    const unsigned long long heads = __ballot(head);
    const unsigned long long above = lane == 63 ? 0ull : (heads | ~act) & (~0ull << (lane + 1));
    const int next = above ? __ffsll(above) - 1 : 64;
    const unsigned long long run = (next == 64 ? ~0ull : ((1ull << next) - 1ull)) & (~0ull << lane) & act;
    const uint32_t n_hi = (uint32_t)__popcll(__ballot(high) & run);
    const uint32_t n_lo = (uint32_t)__popcll(run) - n_hi;
    if (head) {
      if (b32) {
        atomicAdd((uint32_t*)&words[word], n_lo);
      } else {
        const uint32_t old = atomicAdd((uint32_t*)&words[word], n_lo | (n_hi << 16));
        if ((old & 0xFFFFu) + n_lo > 0xFFFFu || (old >> 16) + n_hi > 0xFFFFu)
          atomicOr((uint32_t*)kp.overflow, 1u);
      }
    }
  }
}

template <int BITS, class P>
__device__ __forceinline__ uint32_t load_count(const P* words, uint32_t idx) {
  if (BITS == 32) return words[idx];
  return (words[idx >> 1] >> ((idx & 1u) * 16u)) & 0xFFFFu;
}

// omega_field_update, device_function.cu:373-412: fp64 in the reference's
// operation order, rounded once to fp32.
// With `cadence` set, a system that keeps the reference cadence skips the
// update after sweeps whose count is a multiple of 10 (test.cpp:42-43).
__global__ void __launch_bounds__(BLOCK) k_omega(const KParams* __restrict__ Ps,
                                                 const BlockRef* __restrict__ bt, int cadence) {
  const BlockRef br = block_ref(bt);
  KPC& kp = param(Ps, br.sys);
  if (step_skipped(kp)) return;
  if (cadence && !kp.omega_every_step && ((const uint32_t*)kp.ctr)[0] % 10u == 0u) return;
  const SCMF_GLOBAL uint32_t* __restrict__ words = gptr((const uint32_t*)kp.counts);
  SCMF_GLOBAL float* __restrict__ omega = gptr((float*)kp.omega);
  const uint32_t C = kp.n_cells;
  const int T = kp.T;
  const bool b32 = kp.density_bits == 32;
  const int64_t stride = (int64_t)br.nblk * BLOCK;
  for (int64_t c = (int64_t)br.local * BLOCK + threadIdx.x; c < C; c += stride) {
    double n[MAX_T];
    for (int s = 0; s < T; ++s)
      n[s] = (double)(b32 ? load_count<32>(words, (uint32_t)c + s * C)
                          : load_count<16>(words, (uint32_t)c + s * C));
    for (int t = 0; t < T; ++t) {
      double w = 0;
      for (int s = 0; s < T; ++s) w += n[s] * kp.fst[t];
      w += -1;
      w = kp.inv_nref * (kp.xn[t + T * t] * w);
      for (int s = 0; s < T; ++s) {
        if (t != s) {
          const double dnorm = -0.5 * kp.xn[t + T * s];
          const double interaction = dnorm * kp.inv_nref * (kp.fst[t] * n[t] - kp.fst[s] * n[s]);
          w += interaction;
        }
      }
      omega[c + (int64_t)t * C] = (float)w;
    }
  }
}

// The same update for one system whose cell count is a multiple of four:
// four cells per lane, the counts of each type from one 8- or 16-byte load
// and the omega of each type stored with one 16-byte store.  Each cell's
// arithmetic is k_omega's, in the same order (same bits).
typedef uint32_t u32v2 __attribute__((ext_vector_type(2)));

template <int T, class KP>
__device__ __forceinline__ float omega_one(const KP& kp, const double (&n)[T], int t) {
  double w = 0;
#pragma unroll
  for (int s = 0; s < T; ++s) w += n[s] * kp.fst[t];
  w += -1;
  w = kp.inv_nref * (kp.xn[t + T * t] * w);
#pragma unroll
  for (int s = 0; s < T; ++s) {
    if (t != s) {
      const double dnorm = -0.5 * kp.xn[t + T * s];
      const double interaction = dnorm * kp.inv_nref * (kp.fst[t] * n[t] - kp.fst[s] * n[s]);
      w += interaction;
    }
  }
  return (float)w;
}

template <int T, int BITS>
__global__ void __launch_bounds__(BLOCK) k_omega4(const KParams* __restrict__ Ps, int cadence) {
  KPC& kp = param(Ps, 0);
  if (cadence && !kp.omega_every_step && ((const uint32_t*)kp.ctr)[0] % 10u == 0u) return;
  const SCMF_GLOBAL uint32_t* __restrict__ words = gptr((const uint32_t*)kp.counts);
  SCMF_GLOBAL float* __restrict__ omega = gptr((float*)kp.omega);
  const uint32_t C = kp.n_cells;
  const uint32_t groups = C / 4;
  const uint32_t stride = gridDim.x * BLOCK;
  for (uint32_t gi = blockIdx.x * BLOCK + threadIdx.x; gi < groups; gi += stride) {
    uint32_t raw[T][BITS == 32 ? 4 : 2];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const uint32_t first = (uint32_t)t * C + 4 * gi;  // a multiple of four
      if constexpr (BITS == 32) {
        const u32v4 v = *(const SCMF_GLOBAL u32v4*)(words + first);
        raw[t][0] = v.x;
        raw[t][1] = v.y;
        raw[t][2] = v.z;
        raw[t][3] = v.w;
      } else {
        const u32v2 v = *(const SCMF_GLOBAL u32v2*)(words + first / 2);
        raw[t][0] = v.x;
        raw[t][1] = v.y;
      }
    }
    float out[T][4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double n[T];
#pragma unroll
      for (int t = 0; t < T; ++t)
        n[t] = (double)(BITS == 32 ? raw[t][e] : (raw[t][e >> 1] >> ((e & 1) * 16)) & 0xFFFFu);
#pragma unroll
      for (int t = 0; t < T; ++t) out[t][e] = omega_one<T>(kp, n, t);
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
      f32x4 v = {out[t][0], out[t][1], out[t][2], out[t][3]};
      *(SCMF_GLOBAL f32x4*)(omega + (uint32_t)t * C + 4 * gi) = v;
    }
  }
}

template <int BITS>
const void* omega4_kernel(int T) {
  switch (T) {
    case 1: return (const void*)k_omega4<1, BITS>;
    case 2: return (const void*)k_omega4<2, BITS>;
    case 3: return (const void*)k_omega4<3, BITS>;
    case 4: return (const void*)k_omega4<4, BITS>;
    case 5: return (const void*)k_omega4<5, BITS>;
    case 6: return (const void*)k_omega4<6, BITS>;
    case 7: return (const void*)k_omega4<7, BITS>;
    default: return (const void*)k_omega4<8, BITS>;
  }
}

// Mean squared end-to-end distance: per-block partial sums (summed on the
// host in block order, so the result is reproducible).
__global__ void __launch_bounds__(BLOCK) k_re2(KParams kp, const float* __restrict__ x,
                                               const float* __restrict__ y,
                                               const float* __restrict__ z,
                                               double* __restrict__ partial) {
  __shared__ double red[BLOCK];
  const int64_t chain = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  double v = 0;
  if (chain < kp.n_chains) {
    const int p = arch_of_chain(kp, chain);
    const int64_t slot = kp.chain_slot[chain];
    const int64_t a = kp.bead_base[p] + (slot - kp.chain_base[p]) * kp.len[p];
    const int64_t e = a + kp.len[p] - 1;
    const double ddx = (double)x[a] - (double)x[e];
    const double ddy = (double)y[a] - (double)y[e];
    const double ddz = (double)z[a] - (double)z[e];
    v = ddx * ddx + ddy * ddy + ddz * ddz;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ------------------------------------------------ structure factor
// S(q) = |sum_c w(c) exp(-i q.r_c)|^2 / n_beads with w(c) = sum_t coef[t]
// n_t(c), from an in-place double-precision real-to-complex FFT.

struct SqCoef {
  double c[MAX_T];
};

// w(c) into the padded rows of the in-place transform: row (z, y) holds
// 2*(nx/2+1) doubles, the first nx of them the cells x = 0..nx-1.
__global__ void __launch_bounds__(BLOCK) k_sq_weights(KParams kp, SqCoef coef,
                                                      double* __restrict__ w, int64_t row_pad) {
  const uint32_t* __restrict__ words = kp.counts;
  const uint32_t C = kp.n_cells;
  const bool b32 = kp.density_bits == 32;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x; c < C; c += stride) {
    double v = 0;
    for (int t = 0; t < kp.T; ++t) {
      const uint32_t idx = (uint32_t)c + (uint32_t)t * C;
      const uint32_t n = b32 ? words[idx] : (words[idx >> 1] >> ((idx & 1) * 16)) & 0xFFFFu;
      v += coef.c[t] * (double)n;
    }
    const int64_t row = c / kp.nx;
    w[row * row_pad + (c - row * kp.nx)] = v;
  }
}

constexpr int SQ_LDS_BINS = 2048;

// One lane per stored mode (kz, ky, kx) with 0 <= kx <= nx/2.  Modes with
// 0 < kx < nx/2 (or kx = (nx-1)/2 for odd nx) stand for themselves and
// their conjugate partner at -k, so they count twice.  The shell of a mode
// is the nearest integer to |q|/dq, or to q^2/dq^2 when by_q2 is set.
__global__ void __launch_bounds__(BLOCK) k_sq_shells(const double* __restrict__ F, int nx, int ny,
                                                     int nz, double gx, double gy, double gz,
                                                     double inv_dq, int by_q2, double inv_n,
                                                     int n_bins, double* __restrict__ sum,
                                                     unsigned long long* __restrict__ cnt) {
  __shared__ double lsum[SQ_LDS_BINS];
  __shared__ unsigned lcnt[SQ_LDS_BINS];
  const int lbins = n_bins < SQ_LDS_BINS ? n_bins : SQ_LDS_BINS;
  for (int b = threadIdx.x; b < lbins; b += BLOCK) {
    lsum[b] = 0;
    lcnt[b] = 0;
  }
  __syncthreads();
  const int hx = nx / 2 + 1;
  const int64_t modes = (int64_t)nz * ny * hx;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t m = (int64_t)blockIdx.x * BLOCK + threadIdx.x; m < modes; m += stride) {
    const int64_t r = m / hx;
    const int kx = (int)(m - r * hx);
    const int iy = (int)(r % ny);
    const int iz = (int)(r / ny);
    const int ky = 2 * iy > ny ? iy - ny : iy;
    const int kz = 2 * iz > nz ? iz - nz : iz;
    const double qx = gx * kx, qy = gy * ky, qz = gz * kz;
    const double q2 = qx * qx + qy * qy + qz * qz;
    const double u = by_q2 ? q2 * inv_dq * inv_dq : sqrt(q2) * inv_dq;
    const double ub = rint(u);
    if (!(ub < (double)n_bins)) continue;
    const int b = (int)ub;
    const double re = F[2 * m], im = F[2 * m + 1];
    const unsigned wt = (kx == 0 || 2 * kx == nx) ? 1u : 2u;
    const double v = (double)wt * ((re * re + im * im) * inv_n);
    if (b < SQ_LDS_BINS) {
      atomicAdd(&lsum[b], v);
      atomicAdd(&lcnt[b], wt);
    } else {
      atomicAdd(&sum[b], v);
      atomicAdd(&cnt[b], (unsigned long long)wt);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < lbins; b += BLOCK) {
    if (lcnt[b]) {
      atomicAdd(&sum[b], lsum[b]);
      atomicAdd(&cnt[b], (unsigned long long)lcnt[b]);
    }
  }
}

// Stored index of reference bead r (reference order: chains by id, beads in
// order): the chain's slot replaces its id inside its architecture block.
__device__ __forceinline__ int64_t stored_index(const KParams& kp, int64_t r) {
  const int p = arch_of_bead(kp, r);
  const int len = kp.len[p];
  const int64_t rel = r - kp.bead_base[p];
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
  const int64_t slot = kp.chain_slot[kp.chain_base[p] + q];
  return kp.bead_base[p] + (slot - kp.chain_base[p]) * len + m;
}

// Reference beads [r0, r0 + n) to float [bead][xyz].
__global__ void __launch_bounds__(BLOCK) k_gather_ref(KParams kp, const float* __restrict__ x,
                                                      const float* __restrict__ y,
                                                      const float* __restrict__ z,
                                                      float* __restrict__ aos, int64_t r0,
                                                      int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    const int64_t src = stored_index(kp, r0 + i);
    aos[3 * i] = x[src];
    aos[3 * i + 1] = y[src];
    aos[3 * i + 2] = z[src];
  }
}

__global__ void __launch_bounds__(BLOCK) k_scatter_ref(KParams kp, const float* __restrict__ aos,
                                                       float* __restrict__ x,
                                                       float* __restrict__ y,
                                                       float* __restrict__ z, int64_t r0,
                                                       int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    const int64_t dst = stored_index(kp, r0 + i);
    x[dst] = aos[3 * i];
    y[dst] = aos[3 * i + 1];
    z[dst] = aos[3 * i + 2];
  }
}

__global__ void __launch_bounds__(BLOCK) k_fill_btype(KParams kp, uint8_t* __restrict__ bt) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x; g < kp.n_beads; g += stride) {
    const int p = arch_of_bead(kp, g);
    const int len = kp.len[p];
    const int64_t rel = g - kp.bead_base[p];
    int64_t q = (int64_t)((double)rel * kp.inv_len[p]);
    int64_t m = rel - q * len;
    if (m >= len) m -= len;
    if (m < 0) m += len;
    bt[g] = (uint8_t)kp.monotype[kp.mono_off[p] + m];
  }
}

__global__ void __launch_bounds__(BLOCK) k_iota(uint32_t* __restrict__ a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) a[i] = (uint32_t)i;
}

template <int BITS>
__global__ void __launch_bounds__(BLOCK) k_widen(const uint32_t* __restrict__ words,
                                                 uint32_t* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride)
    out[i] = load_count<BITS>(words, (uint32_t)i);
}

// Staging of scmf_copy_out_noalloc: a fixed device array, no allocation.
// Every GPU the library runs on holds its own copy until the process ends.
constexpr int64_t NOALLOC_WORDS = (int64_t)1 << 22;  // 16 MiB
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

}  // namespace

// ============================================================== host side
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(SCMF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

// The dynamic-LDS limit of a kernel is a process-wide attribute: it is only
// ever raised (to the largest need of any system or batch so far), so a later
// system with a smaller need never lowers it under one already running.
hipError_t raise_lds_limit(const void* fn, size_t bytes) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, size_t>> set;
  std::lock_guard<std::mutex> lock(mu);
  for (auto& e : set)
    if (e.first == fn) {
      if (e.second >= bytes) return hipSuccess;
      const hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
      if (r == hipSuccess) e.second = bytes;
      return r;
    }
  const hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (r == hipSuccess) set.push_back({fn, bytes});
  return r;
}

enum Kind { K_SWEEP = 0, K_GRID = 1, K_OMEGA = 2, K_KINDS = 3 };

// Sweep kernel variants: L lanes per chain (0 = positions in global memory).
struct SweepVariant {
  int L;
  bool wide;  // uint16 bookkeeping (chains longer than 255 beads)
  const void* fn;
  size_t (*bytes)(int nmax);
};

struct Timed {
  int kind;
  hipEvent_t a, b;
};

// Position arrays hold a multiple of four beads (the gridding reads them
// four at a time).
int64_t padded(int64_t beads) { return (beads + 3) & ~(int64_t)3; }

unsigned grid_for(int64_t n) {
  int64_t g = (n + BLOCK - 1) / BLOCK;
  if (g < 1) g = 1;
  if (g > (1 << 30)) g = 1 << 30;
  return (unsigned)g;
}

unsigned grid_strided(int64_t n) {
  int64_t g = (n + BLOCK - 1) / BLOCK;
  if (g < 1) g = 1;
  if (g > 1024 * 64) g = 1024 * 64;
  return (unsigned)g;
}

}  // namespace

struct AsyncOut;

struct scmf_sys {
  int device = 0;
  hipStream_t stream = nullptr;
  HostDesc desc;
  scmf_opts opts{};
  KParams kp{};
  int64_t n_beads = 0, n_chains = 0, n_cells = 0;
  int T = 0;
  int nmax = 0;         // longest chain
  // spatial ordering of the slots and the tiled gridding
  uint32_t* slot_chain = nullptr;
  uint32_t* chain_slot = nullptr;
  uint32_t* slot_chain2 = nullptr;
  uint8_t* btype = nullptr;
  uint32_t* keys = nullptr;
  uint32_t* tcount = nullptr;
  uint32_t* tstart = nullptr;
  uint32_t* tfill = nullptr;
  uint32_t* new_slot = nullptr;
  float* x2 = nullptr;
  float* y2 = nullptr;
  float* z2 = nullptr;
  uint16_t* images = nullptr;  // window images of the tiled gridding
  uint32_t* cells = nullptr;    // packed bead cells written by the sweep (tiled systems)
  uint32_t* cells2 = nullptr;   // the cells' second buffer, for the ordering's permutation
  uint8_t* sched = nullptr;     // move schedules made by the schedule kernels (chains of <= 255 beads)
  bool pre_sched = false;       // the scheduled sweep shapes read them
  bool cells_fresh = false;     // cells match the positions (the last change was such a sweep)
  unsigned long long* d_fallback = nullptr;
  TileGeo geo{};
  size_t tile_lds = 0;  // LDS of k_grid_tiles; 0 = tiled gridding not usable
  bool sorted = false;
  int sort_every = 0;
  uint64_t n_sorts = 0;
  uint64_t n_cell_grids = 0;
  struct Run {
    int64_t c0, c1;             // chains [c0, c1) of consecutive architectures
    int nmax;                   // their longest chain
  };
  std::vector<Run> runs;        // the sweep's launches
  KParams* d_kp = nullptr;  // device copy of kp read by the batch-capable kernels
  uint32_t* d_ctr = nullptr;
  size_t lds_bytes = 0; // LDS of the sweep kernel
  const SweepVariant* variant = nullptr;
  int lds_per_cu = 0, lds_per_block = 0;  // the device's LDS limits
  float* x = nullptr;
  float* y = nullptr;
  float* z = nullptr;
  uint32_t* counts = nullptr;
  int64_t count_words = 0;
  float* omega = nullptr;
  int* d_monotype = nullptr;
  int* d_nbr_start = nullptr;
  int* d_nbr = nullptr;
  unsigned long long* d_accepts = nullptr;
  uint32_t* d_overflow = nullptr;
  double* d_partial = nullptr;
  float* staging = nullptr;
  int64_t staging_beads = 0;
  // structure factor: in-place D2Z plan and buffer, shell accumulators
  hipfftHandle fft = 0;
  bool fft_ready = false;
  double* fft_buf = nullptr;
  double* d_sq_sum = nullptr;
  unsigned long long* d_sq_cnt = nullptr;
  int32_t sq_cap = 0;
  AsyncOut* out = nullptr;  // asynchronous density writer, made on first use
  uint64_t sweeps = 0;
  uint64_t moves = 0;
  bool timing = false;
  std::vector<Timed> pending;
  std::vector<hipEvent_t> pool;
  double t_sum[K_KINDS] = {0, 0, 0};
  uint64_t t_n[K_KINDS] = {0, 0, 0};

  int ev_begin(int kind, Timed* t) {
    t->kind = kind;
    for (hipEvent_t* e : {&t->a, &t->b}) {
      if (!pool.empty()) {
        *e = pool.back();
        pool.pop_back();
      } else {
        HIP_TRY(hipEventCreate(e));
      }
    }
    HIP_TRY(hipEventRecord(t->a, stream));
    return SCMF_OK;
  }
  int ev_end(Timed* t) {
    HIP_TRY(hipEventRecord(t->b, stream));
    pending.push_back(*t);
    return SCMF_OK;
  }
  int drain_timing() {
    if (pending.empty()) return SCMF_OK;
    HIP_TRY(hipStreamSynchronize(stream));
    for (Timed& t : pending) {
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
      t_sum[t.kind] += ms;
      t_n[t.kind] += 1;
      pool.push_back(t.a);
      pool.push_back(t.b);
    }
    pending.clear();
    return SCMF_OK;
  }
};

namespace {

int derive(scmf_sys* s) {
  const HostDesc& d = s->desc;
  KParams& kp = s->kp;
  memset(&kp, 0, sizeof kp);
  kp.T = d.T;
  kp.P = (int)d.poly_length.size();
  kp.nx = d.n[0];
  kp.ny = d.n[1];
  kp.nz = d.n[2];
  kp.Lx = d.L[0];
  kp.Ly = d.L[1];
  kp.Lz = d.L[2];
  kp.dx = kp.Lx / kp.nx;
  kp.dy = kp.Ly / kp.ny;
  kp.dz = kp.Lz / kp.nz;
  kp.ax[0] = make_axis(kp.Lx, kp.nx);
  kp.ax[1] = make_axis(kp.Ly, kp.ny);
  kp.ax[2] = make_axis(kp.Lz, kp.nz);
  kp.pow2 = kp.ax[0].pow2 && kp.ax[1].pow2 && kp.ax[2].pow2;
  s->n_cells = (int64_t)kp.nx * kp.ny * kp.nz;
  s->T = d.T;
  if ((int64_t)d.T * s->n_cells >= (int64_t)0xFFFFFFFF)
    return fail(SCMF_E_ARG, "types x cells must stay below 2^32");
  kp.n_cells = (uint32_t)s->n_cells;
  // init_config.cu:480-484 and :662-664.
  const double cste = 1.0 / sqrt(3.0 * ((double)d.nref - 1.0));
  kp.harm = 1.0 / (2.0 * cste * cste);
  kp.inv_nref = 1 / (double)d.nref;
  for (int t = 0; t < d.T; ++t)
    for (int u = 0; u < d.T; ++u) kp.xn[t + d.T * u] = d.xn[t + (size_t)d.T * u];
  int64_t cb = 0, bb = 0;
  int mo = 0;
  for (int p = 0; p < kp.P; ++p) {
    kp.chain_base[p] = cb;
    kp.bead_base[p] = bb;
    kp.mono_off[p] = mo;
    kp.len[p] = d.poly_length[p];
    kp.inv_len[p] = 1.0 / (double)d.poly_length[p];
    s->nmax = std::max(s->nmax, (int)d.poly_length[p]);
    cb += d.n_polymers[p];
    bb += d.n_polymers[p] * (int64_t)d.poly_length[p];
    mo += d.poly_length[p];
  }
  kp.chain_base[kp.P] = cb;
  kp.bead_base[kp.P] = bb;
  kp.mono_off[kp.P] = mo;
  s->n_chains = cb;
  s->n_beads = bb;
  kp.n_chains = cb;
  kp.n_beads = bb;
  // init_config.cu:784-785 and :804-809.
  for (int t = 0; t < d.T; ++t) {
    kp.fst[t] = (double)s->n_cells / (double)s->n_beads;
    const double A = d.dtNxi / d.nref;
    kp.A[t] = (float)A;
    kp.R[t] = (float)sqrt(A * 2);
  }
  kp.stop = NO_STOP;
  kp.seed = s->opts.seed;
  kp.system_id = s->opts.system_id;
  kp.move_type = s->opts.move_type;
  kp.trial_delta = s->opts.trial_delta;
  return SCMF_OK;
}

int upload_arch(scmf_sys* s) {
  const HostDesc& d = s->desc;
  std::vector<int> nbr_start, nbr;
  size_t coff = 0;
  for (size_t p = 0; p < d.poly_length.size(); ++p) {
    const int N = d.poly_length[p];
    bool linear = true;
    for (int i = 0; i < N; ++i) {
      nbr_start.push_back((int)nbr.size());
      int deg = 0;
      for (int j = 0; j < N; ++j) {
        const bool bonded = d.connection[coff + (size_t)i * N + j] != 0;
        if (bonded) {
          nbr.push_back(j);
          ++deg;
        }
        if (bonded != (j == i - 1 || j == i + 1)) linear = false;
      }
      if (deg > MAX_DEG) return fail(SCMF_E_ARG, "a bead has more than 8 bonds");
    }
    s->kp.linear[p] = linear ? 1 : 0;
    coff += (size_t)N * N;
  }
  // types as register masks when every chain fits 64 bits of one-bit types
  bool bits = d.T <= 2;
  for (size_t p = 0; p < d.poly_length.size(); ++p) bits = bits && d.poly_length[p] <= 64;
  s->kp.type_bits = bits ? 1 : 0;
  for (size_t p = 0, m = 0; p < d.poly_length.size(); ++p) {
    uint64_t tb = 0;
    for (int b = 0; b < d.poly_length[p]; ++b, ++m)
      if (bits && d.monotype[m] != 0) tb |= (uint64_t)1 << b;
    s->kp.tbits[p] = tb;
  }
  nbr_start.push_back((int)nbr.size());
  if (nbr.empty()) nbr.push_back(0);
  std::vector<int> mono(d.monotype.begin(), d.monotype.end());
  HIP_TRY(hipMalloc(&s->d_monotype, sizeof(int) * mono.size()));
  HIP_TRY(hipMalloc(&s->d_nbr_start, sizeof(int) * nbr_start.size()));
  HIP_TRY(hipMalloc(&s->d_nbr, sizeof(int) * nbr.size()));
  HIP_TRY(hipMemcpy(s->d_monotype, mono.data(), sizeof(int) * mono.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_nbr_start, nbr_start.data(), sizeof(int) * nbr_start.size(),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s->d_nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice));
  s->kp.monotype = s->d_monotype;
  s->kp.nbr_start = s->d_nbr_start;
  s->kp.nbr = s->d_nbr;
  return SCMF_OK;
}

int ensure_staging(scmf_sys* s, int64_t beads) {
  if (s->staging_beads >= beads) return SCMF_OK;
  if (s->staging) HIP_TRY(hipFree(s->staging));
  s->staging = nullptr;
  s->staging_beads = 0;
  HIP_TRY(hipMalloc(&s->staging, sizeof(float) * 3 * beads));
  s->staging_beads = beads;
  return SCMF_OK;
}

constexpr int64_t STAGE_BEADS = (int64_t)1 << 24;

// Refresh the device copy of the system parameters (after creation, a
// reordering, or a re-keyed stream).
int sync_kp(scmf_sys* s) {
  HIP_TRY(hipMemcpyAsync(s->d_kp, &s->kp, sizeof(KParams), hipMemcpyHostToDevice, s->stream));
  return SCMF_OK;
}

// Set the device sweep word (the stream position of the next sweep).
int set_ctr(scmf_sys* s, uint32_t next_sweep) {
  const uint32_t v[2] = {next_sweep, 0u};
  HIP_TRY(hipMemcpyAsync(s->d_ctr, v, sizeof v, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

size_t lds1_bytes(int nmax) { return (size_t)3 * nmax * 64 * sizeof(float); }
size_t none_bytes(int) { return 0; }

const SweepVariant kVariants[] = {
    {0, false, (const void*)k_sweep_global, none_bytes},
    {1, false, (const void*)k_sweep_lds, lds1_bytes},
    {2, false, (const void*)k_sweep_sched<2, uint8_t>, sched_lds_bytes<2, uint8_t>},
    {4, false, (const void*)k_sweep_sched<4, uint8_t>, sched_lds_bytes<4, uint8_t>},
    {8, false, (const void*)k_sweep_sched<8, uint8_t>, sched_lds_bytes<8, uint8_t>},
    {2, true, (const void*)k_sweep_sched<2, uint16_t>, sched_lds_bytes<2, uint16_t>},
    {4, true, (const void*)k_sweep_sched<4, uint16_t>, sched_lds_bytes<4, uint16_t>},
    {8, true, (const void*)k_sweep_sched<8, uint16_t>, sched_lds_bytes<8, uint16_t>},
    // 16 lanes a chain, 4 chains a wave: twice the waves of 8 lanes for small
    // systems of short chains, whose rounds hardly grow (N = 32: 6.0 against
    // 6.1 rounds a sweep)
    {16, false, (const void*)k_sweep_sched<16, uint8_t>, sched_lds_bytes<16, uint8_t>},
    {16, true, (const void*)k_sweep_sched<16, uint16_t>, sched_lds_bytes<16, uint16_t>},
};

// Pick the variant: the requested lanes_per_chain if it fits, else the
// smallest L that keeps at least 8 wavefronts per CU resident by LDS, else
// the one with the most resident wavefronts.
int choose_sweep(scmf_sys* s) {
  int max_block = 0, per_cu = 0;
  (void)hipDeviceGetAttribute(&max_block, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device);
  (void)hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                              s->device);
  if (per_cu <= 0) per_cu = max_block;
  s->lds_per_cu = per_cu;
  s->lds_per_block = max_block;
  const bool wide = s->nmax > 255;
  const SweepVariant* best = &kVariants[0];
  int best_waves = 0;
  if (s->opts.lanes_per_chain == -1) {  // positions in global memory, one lane per chain
    s->variant = &kVariants[0];
    s->lds_bytes = 0;
    return SCMF_OK;
  }
  for (const SweepVariant& v : kVariants) {
    if (v.L == 0 || (v.L >= 2 && v.wide != wide)) continue;
    const size_t need = v.bytes(s->nmax);
    if (need > (size_t)max_block) continue;
    const int waves = need ? (int)(per_cu / need) : 64;
    if (s->opts.lanes_per_chain == v.L) {
      best = &v;
      break;
    }
    if (s->opts.lanes_per_chain == 0) {
      const bool good = waves >= 8, best_good = best_waves >= 8;
      if (best->L == 0 || (good && !best_good) || (!good && !best_good && waves > best_waves)) {
        best = &v;
        best_waves = waves;
      }
    }
  }
  if (s->opts.lanes_per_chain != 0 && best->L != s->opts.lanes_per_chain)
    return fail(SCMF_E_ARG, "lanes_per_chain must be 1, 2, 4, 8 or 16 and fit in LDS");
  const size_t need = best->bytes(s->nmax);
  if (need > 65536) {
    hipError_t e = raise_lds_limit(best->fn, need);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      best = &kVariants[0];
    }
  }
  s->variant = best;
  s->lds_bytes = s->pre_sched && best->L >= 2 ? sched_pre_lds_bytes(best->L, s->nmax)
                                              : best->bytes(s->nmax);
  return SCMF_OK;
}

// k_schedule_reg when its register counts hold the rounds, else k_schedule_lds.
// The schedules of the chains in [c0, c1), whose longest has nmax beads, by
// the register-count kernel, which stores only its own chains' record bytes;
// when that kernel cannot take them, the LDS kernel schedules every chain of
// the system (its 16-byte stores write the records of whole 64-chain blocks),
// so the caller then launches it once.
int launch_schedule(scmf_sys* s, int cap, int64_t c0, int64_t c1, int nmax) {
  const KParams* kp = s->d_kp;
  if (schedule_in_registers(nmax, cap)) {
    const dim3 grid((unsigned)((c1 - c0 + SCHED_LANES - 1) / SCHED_LANES));
    const size_t lds = schedule_lds_bytes(nmax, cap);
    const bool two = nmax > 64;
    const void* fn = cap == 2 ? (two ? (const void*)k_schedule_reg<1, 2> : (const void*)k_schedule_reg<1, 1>)
                   : cap == 4 ? (two ? (const void*)k_schedule_reg<2, 2> : (const void*)k_schedule_reg<2, 1>)
                   : cap == 8 ? (two ? (const void*)k_schedule_reg<3, 2> : (const void*)k_schedule_reg<3, 1>)
                              : (two ? (const void*)k_schedule_reg<4, 2> : (const void*)k_schedule_reg<4, 1>);
    void* args[] = {(void*)&kp, (void*)&nmax, (void*)&c0, (void*)&c1};
    HIP_TRY(hipLaunchKernel(fn, grid, dim3(SCHED_LANES), args, lds, s->stream));
  } else {
    const dim3 grid((unsigned)((s->n_chains + SCHED_LANES - 1) / SCHED_LANES));
    const size_t lds = schedule_lds_bytes(s->nmax, cap);
    if (lds > 65536) HIP_TRY(raise_lds_limit((const void*)k_schedule_lds, lds));
    void* args[] = {(void*)&kp, (void*)&s->nmax, (void*)&cap};
    HIP_TRY(hipLaunchKernel((const void*)k_schedule_lds, grid, dim3(SCHED_LANES), args, lds,
                            s->stream));
  }
  return SCMF_OK;
}

// The variant that sweeps a run of chains of at most nmax beads: the chosen
// one, except that at 8 lanes a chain a run of chains longer than 64 beads
// (whose LDS tiles would leave fewer waves resident) takes 16 lanes a chain
// with schedules in rounds of 16 when that keeps more waves resident (the
// 128-bead run of the 1.0e9-bead blend: 2.5 waves a SIMD at 8 lanes, 5 at 16).
const SweepVariant* run_variant(const scmf_sys* s, const SweepVariant* v, int nmax) {
  if (v->L != 8 || !s->pre_sched || nmax <= 64 || s->lds_per_cu <= 0) return v;
  const SweepVariant* v16 = nullptr;
  for (const SweepVariant& w : kVariants)
    if (w.L == 16 && w.wide == v->wide) v16 = &w;
  if (!v16) return v;
  auto resident = [&](int L) {  // waves per CU that LDS allows, at most the 20 registers allow
    return std::min(20, (int)((size_t)s->lds_per_cu / sched_pre_lds_bytes(L, nmax)));
  };
  return resident(16) > resident(8) && sched_pre_lds_bytes(16, nmax) <= (size_t)s->lds_per_block ? v16 : v;
}

// clear: the sweep also clears the counts (scmf_step, whose gridding follows);
// advance: k_advance moves the stream sweep word on (else the step's gridding does).
// The LDS shapes sweep each run of architectures (s->runs) with its own
// launch, the chain tiles sized to the run's longest chain, so a few long
// chains do not cut the waves of all the others.
int launch_sweep(scmf_sys* s, bool clear = false, bool advance = true) {
  const SweepVariant* v = s->variant;
  const BlockRef* none = nullptr;
  const KParams* kpp = s->d_kp;
  if (v->L == 0) {
    int clr = clear ? 1 : 0;
    int64_t c0 = 0, c1 = s->n_chains;
    hipLaunchKernelGGL(k_sweep_global, dim3(grid_for(s->n_chains)), dim3(BLOCK), 0, s->stream,
                       kpp, none, clr, c0, c1);
  } else {
    // each run's variant (and so its schedules' round size)
    std::vector<const SweepVariant*> vr(s->runs.size(), v);
    if (s->pre_sched) {
      bool split = true;
      for (size_t k = 0; k < s->runs.size(); ++k) {
        vr[k] = run_variant(s, v, s->runs[k].nmax);
        split = split && schedule_in_registers(s->runs[k].nmax, vr[k]->L);
      }
      if (split) {
        for (size_t k = 0; k < s->runs.size(); ++k) {
          const auto& r = s->runs[k];
          const int rc = launch_schedule(s, vr[k]->L, r.c0, r.c1, r.nmax);
          if (rc != SCMF_OK) return rc;
        }
      } else {  // one schedule launch over every chain: one round size for all runs
        std::fill(vr.begin(), vr.end(), v);
        const int rc = launch_schedule(s, v->L, 0, s->n_chains, s->nmax);
        if (rc != SCMF_OK) return rc;
      }
    }
    for (size_t k = 0; k < s->runs.size(); ++k) {
      const SweepVariant* w = vr[k];
      const int G = 64 / w->L;
      int nm = s->runs[k].nmax;
      int clr = clear && k == 0 ? 1 : 0;  // the first launch clears every count
      int64_t c0 = s->runs[k].c0, c1 = s->runs[k].c1;
      const size_t lds = s->pre_sched && w->L >= 2 ? sched_pre_lds_bytes(w->L, nm) : w->bytes(nm);
      if (lds > 65536) HIP_TRY(raise_lds_limit(w->fn, lds));
      const dim3 grid((unsigned)((c1 - c0 + G - 1) / G));
      void* args[] = {(void*)&kpp, (void*)&none, (void*)&nm, (void*)&clr, (void*)&c0, (void*)&c1};
      HIP_TRY(hipLaunchKernel(w->fn, grid, dim3(64), args, lds, s->stream));
    }
  }
  if (advance)
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s->stream, (const KParams*)s->d_kp, 1);
  HIP_TRY(hipGetLastError());
  return SCMF_OK;
}

// Tile geometry of the tiled gridding.  The halo covers the reach of a
// typical chain from its middle bead (0.75 Re scaled by sqrt((N-1)/(Nref-1))
// for the bead-weighted median length N); beads beyond it are counted with
// global atomics.  Among the tilings whose LDS window fits, the one with the
// least window area per interior cell wins (that ratio sets the image
// traffic), with a penalty for one workgroup per CU and for too few tiles to
// fill the GPU.
size_t tile_lds_cap() { return (size_t)160 << 10; }

// The plan depends only on the grid, the cell spacing, the chain lengths
// (with their bead counts), Nref and the number of monomer types, so the
// device-memory estimate of a description uses it too.  Returns the LDS
// bytes of k_grid_tiles and fills *out, or returns 0 if no tiling fits.
size_t plan_tiles(const int n[3], const double d[3], std::vector<std::pair<int, int64_t>> lens,
                  int nref, int T, TileGeo* out) {
  if (lens.empty()) return 0;
  int64_t total = 0;
  for (auto& l : lens) total += l.second;
  std::sort(lens.begin(), lens.end());
  int nmed = lens.back().first;
  int64_t acc = 0;
  for (auto& l : lens) {
    acc += l.second;
    if (2 * acc >= total) {
      nmed = l.first;
      break;
    }
  }
  const double reach = 0.75 * sqrt((double)std::max(nmed - 1, 1) / (double)std::max(nref - 1, 1));
  int hh[3];
  for (int a = 0; a < 3; ++a) hh[a] = std::max(1, std::min(16, (int)ceil(reach / d[a])));
  // LDS: the window, the per-axis offset tables, the type table, slack
  const size_t tables = (size_t)4 * (n[0] + n[1] + n[2]);
  if (tables > ((size_t)16 << 10)) return 0;
  const size_t lds_cap = std::min(((size_t)160 << 10) - MONO_LDS - 512 - tables, tile_lds_cap());
  // candidate tile counts per axis: one tile, or pieces no shorter than
  // twice the halo and no longer than 64 cells
  std::vector<int> cand[3];
  for (int a = 0; a < 3; ++a) {
    cand[a].push_back(1);
    for (int nt = 2; nt <= n[a]; ++nt) {
      if (n[a] / nt < 2 * hh[a]) break;
      if ((n[a] + nt - 1) / nt <= 64) cand[a].push_back(nt);
    }
  }
  double best = 0;
  TileGeo bg{};
  for (int n0 : cand[0])
    for (int n1 : cand[1])
      for (int n2 : cand[2]) {
        const int nts[3] = {n0, n1, n2};
        TileGeo g{};
        int64_t wc = 1, ic = 1;
        for (int a = 0; a < 3; ++a) {
          g.nt[a] = nts[a];
          g.q[a] = n[a] / nts[a];
          g.rem[a] = n[a] % nts[a];
          g.h[a] = nts[a] == 1 ? 0 : hh[a];
          g.E[a] = g.q[a] + (g.rem[a] ? 1 : 0);
          g.W[a] = g.E[a] + 2 * g.h[a];
          wc *= g.W[a];
          ic *= g.E[a];
        }
        const size_t bytes = (size_t)((((wc * T + 1) / 2) + 3) & ~(int64_t)3) * 4;
        if (bytes > lds_cap) continue;
        const int64_t ntiles = (int64_t)n0 * n1 * n2;
        const int per_cu =
            (int)std::min<size_t>(2, ((size_t)160 << 10) / (bytes + tables + MONO_LDS + 512));
        double cost = (double)wc / (double)ic;
        if (per_cu < 2) cost *= 1.15;
        if (ntiles < 1024) cost *= 1024.0 / (double)ntiles;
        const double score = 1.0 / cost;
        if (score > best) {
          best = score;
          bg = g;
          bg.ntiles = (int)ntiles;
          bg.win_cells = (int)wc;
          bg.win_words = (int)((wc * T + 1) / 2);
          bg.win_words_padded = (bg.win_words + 3) & ~3;
        }
      }
  if (best <= 0) return 0;
  // region offsets of a window image: each region sized for the largest tile
  int off = 0;
  for (int r = 0; r < 27; ++r) {
    const int rr[3] = {r % 3, (r / 3) % 3, r / 9};
    int64_t e = T;
    for (int a = 0; a < 3; ++a) e *= rr[a] == 1 ? bg.E[a] : bg.h[a];
    bg.reg_off[r] = off;
    off += (int)((e + 1) & ~(int64_t)1);
  }
  // (two monomer types store the window itself: win_words_padded words)
  if (T == 2) off = std::max(off, 2 * bg.win_words_padded);
  bg.slot_elems = (off + 63) & ~63;
  *out = bg;
  return (size_t)bg.win_words_padded * 4 + tables;
}

void choose_tiles(scmf_sys* s) {
  const KParams& kp = s->kp;
  std::vector<std::pair<int, int64_t>> lens;
  for (int p = 0; p < kp.P; ++p) lens.push_back({kp.len[p], kp.bead_base[p + 1] - kp.bead_base[p]});
  const int n[3] = {kp.nx, kp.ny, kp.nz};
  const double d[3] = {kp.dx, kp.dy, kp.dz};
  s->tile_lds = plan_tiles(n, d, lens, s->desc.nref, s->T, &s->geo);
}

// The tiled gridding kernel: from the sweep's cells or from positions,
// 16- or 32-bit counts, power-of-two cells (positions only), the paired
// window layout of two monomer types.
const void* grid_tiles_kernel(bool cells, bool b32, bool p2, bool pair) {
  if (pair) {
    if (cells) return b32 ? (const void*)k_grid_tiles<32, false, true, true>
                          : (const void*)k_grid_tiles<16, false, true, true>;
    if (b32) return p2 ? (const void*)k_grid_tiles<32, true, false, true>
                       : (const void*)k_grid_tiles<32, false, false, true>;
    return p2 ? (const void*)k_grid_tiles<16, true, false, true>
              : (const void*)k_grid_tiles<16, false, false, true>;
  }
  if (cells) return b32 ? (const void*)k_grid_tiles<32, false, true, false>
                        : (const void*)k_grid_tiles<16, false, true, false>;
  if (b32) return p2 ? (const void*)k_grid_tiles<32, true, false, false>
                     : (const void*)k_grid_tiles<32, false, false, false>;
  return p2 ? (const void*)k_grid_tiles<16, true, false, false>
            : (const void*)k_grid_tiles<16, false, false, false>;
}

int alloc_sort_buffers(scmf_sys* s) {
  if (s->keys) return SCMF_OK;
  const int64_t C = s->n_chains, B = s->n_beads;
  const int64_t nk = (int64_t)s->kp.P * s->geo.ntiles;
  auto A = [&](void** p, size_t bytes) -> int {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SCMF_E_NOMEM, "device allocation of " + std::to_string(bytes) +
                                    " bytes for the spatial ordering failed");
    }
    return SCMF_OK;
  };
  int rc;
  if ((rc = A((void**)&s->keys, sizeof(uint32_t) * C)) ||
      (rc = A((void**)&s->new_slot, sizeof(uint32_t) * C)) ||
      (rc = A((void**)&s->slot_chain2, sizeof(uint32_t) * C)) ||
      (rc = A((void**)&s->tcount, sizeof(uint32_t) * nk)) ||
      (rc = A((void**)&s->tfill, sizeof(uint32_t) * nk)) ||
      (rc = A((void**)&s->tstart, sizeof(uint32_t) * (nk + 1))) ||
      (rc = A((void**)&s->x2, sizeof(float) * padded(B))) ||
      (rc = A((void**)&s->y2, sizeof(float) * padded(B))) ||
      (rc = A((void**)&s->z2, sizeof(float) * padded(B))) ||
      (rc = A((void**)&s->images,
              sizeof(uint16_t) * ((size_t)s->geo.ntiles * s->geo.slot_elems + 64))))
    return rc;
  // the sweep stores bead cells for the next gridding where 10 bits per axis hold them
  if (s->kp.nx <= CELL_AXIS_MAX && s->kp.ny <= CELL_AXIS_MAX && s->kp.nz <= CELL_AXIS_MAX) {
    if ((rc = A((void**)&s->cells, sizeof(uint32_t) * padded(B))) ||
        (rc = A((void**)&s->cells2, sizeof(uint32_t) * padded(B))))
      return rc;
    HIP_TRY(hipMemsetAsync(s->cells, 0, sizeof(uint32_t) * padded(B), s->stream));
    s->kp.cells = s->cells;
    s->cells_fresh = false;
  }
  // the window sums read absent contributions from the element after the
  // last image, which nothing writes
  HIP_TRY(hipMemsetAsync(s->images + (size_t)s->geo.ntiles * s->geo.slot_elems, 0,
                         sizeof(uint16_t) * 64, s->stream));
  if (s->tile_lds + MONO_LDS + 128 > 65536) {
    for (int v = 0; v < 16; ++v) {
      const void* f = grid_tiles_kernel(v & 1, v & 2, v & 4, v & 8);
      if (raise_lds_limit(f, s->tile_lds) !=
          hipSuccess) {
        (void)hipGetLastError();
        return fail(SCMF_E_HIP, "cannot raise the LDS limit of the tiled gridding");
      }
    }
  }
  return SCMF_OK;
}

// Reorder the slots by (architecture, tile of the middle bead).
int sort_chains(scmf_sys* s) {
  if (!s->tile_lds) return SCMF_OK;
  int rc = alloc_sort_buffers(s);
  if (rc) return rc;
  const int64_t C = s->n_chains;
  const int nk = s->kp.P * s->geo.ntiles;
  hipLaunchKernelGGL(k_zero, dim3(grid_strided(nk)), dim3(BLOCK), 0, s->stream, s->tcount,
                     (int64_t)nk);
  hipLaunchKernelGGL(k_tile_keys, dim3(grid_for(C)), dim3(256), 0, s->stream, s->kp, s->geo, s->x,
                     s->y, s->z, s->keys, s->tcount);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s->stream, s->tcount, s->tstart,
                     s->tfill, nk);
  hipLaunchKernelGGL(k_tile_scatter, dim3(grid_for(C)), dim3(256), 0, s->stream, C, s->keys,
                     s->tstart, s->tfill, s->new_slot);
  // cells that match the positions move with them (a bead's cell does not
  // depend on its slot), so the next gridding still reads cells
  const bool carry = s->cells_fresh && s->cells2 != nullptr;
  hipLaunchKernelGGL(k_permute, dim3(grid_strided(s->n_beads)), dim3(256), 0, s->stream, s->kp,
                     s->x, s->y, s->z, s->x2, s->y2, s->z2, carry ? s->cells : nullptr,
                     carry ? s->cells2 : nullptr, s->new_slot, s->slot_chain, s->slot_chain2,
                     s->chain_slot);
  HIP_TRY(hipGetLastError());
  std::swap(s->x, s->x2);
  std::swap(s->y, s->y2);
  std::swap(s->z, s->z2);
  std::swap(s->slot_chain, s->slot_chain2);
  if (carry) {
    std::swap(s->cells, s->cells2);
    s->kp.cells = s->cells;
  }
  s->kp.slot_chain = s->slot_chain;
  s->kp.x = s->x;
  s->kp.y = s->y;
  s->kp.z = s->z;
  s->cells_fresh = carry;
  if ((rc = sync_kp(s))) return rc;
  s->sorted = true;
  s->n_sorts += 1;
  return SCMF_OK;
}

// The tiled gridding of one system whose parameters sit at Ps[sys] (the
// counts must already be cleared).
// use_cells: read the bead cells the last sweep stored (valid only while
// s->cells_fresh holds) instead of binning the positions.
int launch_tiled_grid(scmf_sys* s, hipStream_t st, const KParams* Ps, int sys, bool use_cells,
                      int advance) {
  const dim3 tiles((unsigned)s->geo.ntiles);
  const void* bin = grid_tiles_kernel(use_cells, s->opts.density_bits == 32, s->kp.pow2 != 0,
                                      s->T == 2);
  const uint32_t* tstart = s->tstart;
  void* args[] = {(void*)&Ps, (void*)&sys, (void*)&s->geo, (void*)&tstart, (void*)&s->images,
                  (void*)&s->d_fallback};
  HIP_TRY(hipLaunchKernel(bin, tiles, dim3(GRID_THREADS), args, s->tile_lds, st));
  if (s->T == 2) {
    auto sum = s->opts.density_bits == 32 ? k_tiles_sum_pair<32> : k_tiles_sum_pair<16>;
    hipLaunchKernelGGL(sum, tiles, dim3(SUM_THREADS), 0, st, Ps, sys, s->geo,
                       (const uint32_t*)s->images, advance);
    HIP_TRY(hipGetLastError());
    return SCMF_OK;
  }
  auto sum = s->opts.density_bits == 32 ? k_tiles_sum<32> : k_tiles_sum<16>;
  hipLaunchKernelGGL(sum, tiles, dim3(SUM_THREADS), 0, st, Ps, sys, s->geo,
                     (const uint16_t*)s->images, advance);
  HIP_TRY(hipGetLastError());
  return SCMF_OK;
}

// n sweeps; fused: the first clears the counts and leaves the stream word to
// the step's gridding (scmf_step, n = 1).
int sweeps_impl(scmf_sys* s, int32_t n_sweeps, bool fused) {
  for (int32_t i = 0; i < n_sweeps; ++i) {
    Timed t;
    if (s->timing) {
      int rc = s->ev_begin(K_SWEEP, &t);
      if (rc) return rc;
    }
    int rc = launch_sweep(s, fused, !fused);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    s->cells_fresh = s->kp.cells != nullptr && s->variant->L >= 2;
    if (s->timing) {
      int rc = s->ev_end(&t);
      if (rc) return rc;
    }
    s->sweeps += 1;
    s->moves += (uint64_t)s->n_beads;
  }
  return SCMF_OK;
}

// The gridding; clear: clear the counts first (else the step's sweep did);
// advance: move the stream sweep word on (the step's sweep left it).
int grid_impl(scmf_sys* s, bool clear, bool advance) {
  const int adv = advance ? 1 : 0;
  if (s->tile_lds && !s->sorted && s->sort_every > 0) {
    int rc = sort_chains(s);
    if (rc) return rc;
  }
  Timed t;
  if (s->timing) {
    int rc = s->ev_begin(K_GRID, &t);
    if (rc) return rc;
  }
  if (clear)
    hipLaunchKernelGGL(k_zero_counts, dim3(grid_strided(s->count_words)), dim3(BLOCK), 0, s->stream,
                       (const KParams*)s->d_kp, (const BlockRef*)nullptr);
  const bool b32 = s->opts.density_bits == 32;
  if (s->tile_lds && s->sorted) {
    if (s->cells_fresh) s->n_cell_grids += 1;
    int rc = launch_tiled_grid(s, s->stream, s->d_kp, 0, s->cells_fresh, adv);
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(k_grid, dim3(grid_strided(s->n_beads)), dim3(BLOCK), 0, s->stream,
                       (const KParams*)s->d_kp, (const BlockRef*)nullptr, adv);
  }
  HIP_TRY(hipGetLastError());
  if (s->timing) return s->ev_end(&t);
  return SCMF_OK;
}

int free_all(scmf_sys* s) {
  for (void* p : {(void*)s->x, (void*)s->y, (void*)s->z, (void*)s->counts, (void*)s->omega,
                  (void*)s->d_monotype, (void*)s->d_nbr_start, (void*)s->d_nbr,
                  (void*)s->d_accepts, (void*)s->d_overflow, (void*)s->d_partial,
                  (void*)s->staging, (void*)s->slot_chain, (void*)s->chain_slot,
                  (void*)s->slot_chain2, (void*)s->btype, (void*)s->keys, (void*)s->tcount,
                  (void*)s->tstart, (void*)s->tfill, (void*)s->new_slot, (void*)s->x2,
                  (void*)s->y2, (void*)s->z2, (void*)s->images, (void*)s->d_fallback,
                  (void*)s->d_kp, (void*)s->d_ctr, (void*)s->fft_buf, (void*)s->d_sq_sum,
                  (void*)s->d_sq_cnt, (void*)s->cells, (void*)s->cells2, (void*)s->sched})
    if (p) (void)hipFree(p);
  if (s->fft_ready) (void)hipfftDestroy(s->fft);
  for (Timed& t : s->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t e : s->pool) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  return SCMF_OK;
}

int create_impl(scmf_sys** out, const HostDesc& hd, const scmf_opts* opts) {
  if (!out) return fail(SCMF_E_ARG, "out is null");
  *out = nullptr;
  scmf_desc view = hd.view();
  std::string v = validate_desc(&view);
  if (!v.empty()) return fail(SCMF_E_ARG, v);
  if (hd.T > MAX_T) return fail(SCMF_E_ARG, "at most 8 monomer types");
  if ((int)hd.poly_length.size() > MAX_P) return fail(SCMF_E_ARG, "at most 16 polymer types");
  scmf_opts o{};
  if (opts) o = *opts;
  else {
    o.move_type = SCMF_MOVE_SMART;
    o.density_bits = 32;
    o.init_coords = 1;
  }
  if (o.density_bits != 16 && o.density_bits != 32)
    return fail(SCMF_E_ARG, "density_bits must be 16 or 32");
  if (o.move_type != SCMF_MOVE_SMART && o.move_type != SCMF_MOVE_TRIAL)
    return fail(SCMF_E_ARG, "unknown move type");
  if (o.move_type == SCMF_MOVE_TRIAL && !(o.trial_delta > 0))
    return fail(SCMF_E_ARG, "trial moves need trial_delta > 0");
  for (int32_t N : hd.poly_length)
    if (o.init_coords && N > MAX_INIT_N)
      return fail(SCMF_E_ARG, "initial walk supports chains up to 4096 beads");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SCMF_E_NODEV, "no HIP device available");
  scmf_sys* s = new scmf_sys();
  s->desc = hd;
  s->opts = o;
  int rc = derive(s);
  auto bail = [&](int code) {
    free_all(s);
    delete s;
    return code;
  };
  if (rc) return bail(rc);
  if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(SCMF_E_HIP, "hipGetDevice failed"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SCMF_E_HIP, "hipStreamCreate failed"));
  if ((rc = upload_arch(s))) return bail(rc);
  // Schedules in their own kernel where the in-sweep scan outweighs a launch
  // (measured: slower for systems of a few thousand chains, batches included).
  s->pre_sched = s->nmax <= 255 && s->n_chains >= 65536;
  // sweep runs: consecutive architectures of one chain length share a launch
  // sized to that length (large systems; small ones keep one launch)
  s->runs.clear();
  for (int p = 0; p < s->kp.P; ++p) {
    const int64_t c0 = s->kp.chain_base[p], c1 = s->kp.chain_base[p + 1];
    if (c1 <= c0) continue;
    const int len = s->kp.len[p];
    if (!s->runs.empty() && (s->runs.back().nmax == len || s->n_chains < 65536)) {
      s->runs.back().c1 = c1;
      s->runs.back().nmax = std::max(s->runs.back().nmax, len);
    } else {
      s->runs.push_back({c0, c1, len});
    }
  }
  if ((rc = choose_sweep(s))) return bail(rc);
  const int64_t B = s->n_beads;
  const int64_t TC = (int64_t)s->T * s->n_cells;
  s->count_words = o.density_bits == 32 ? TC : (TC + 1) / 2;
  auto alloc = [&](void** p, size_t bytes) -> int {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SCMF_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return SCMF_OK;
  };
  if ((rc = alloc((void**)&s->x, sizeof(float) * padded(B))) ||
      (rc = alloc((void**)&s->y, sizeof(float) * padded(B))) ||
      (rc = alloc((void**)&s->z, sizeof(float) * padded(B))) ||
      (rc = alloc((void**)&s->counts, sizeof(uint32_t) * s->count_words)) ||
      (rc = alloc((void**)&s->omega, sizeof(float) * TC)) ||
      (rc = alloc((void**)&s->d_accepts, sizeof(unsigned long long) * ACC_SLOTS)) ||
      (rc = alloc((void**)&s->d_overflow, sizeof(uint32_t))) ||
      (rc = alloc((void**)&s->d_partial, sizeof(double) * grid_for(s->n_chains))))
    return bail(rc);
  if ((rc = alloc((void**)&s->slot_chain, sizeof(uint32_t) * s->n_chains)) ||
      (rc = alloc((void**)&s->chain_slot, sizeof(uint32_t) * s->n_chains)) ||
      (rc = alloc((void**)&s->btype, (size_t)B)) ||
      (rc = alloc((void**)&s->d_fallback, sizeof(unsigned long long))))
    return bail(rc);
  if ((rc = alloc((void**)&s->d_kp, sizeof(KParams))) ||
      (rc = alloc((void**)&s->d_ctr, 2 * sizeof(uint32_t))))
    return bail(rc);
  if (s->pre_sched) {
    s->kp.sched_stride = 3 * s->nmax + 3;
    if ((rc = alloc((void**)&s->sched, schedule_record_bytes(s->nmax, s->n_chains))))
      return bail(rc);
    s->kp.sched = s->sched;
  }
  s->kp.slot_chain = s->slot_chain;
  s->kp.chain_slot = s->chain_slot;
  s->kp.btype = s->btype;
  s->kp.x = s->x;
  s->kp.y = s->y;
  s->kp.z = s->z;
  s->kp.omega = s->omega;
  s->kp.counts = s->counts;
  s->kp.count_words = s->count_words;
  s->kp.density_bits = o.density_bits;
  s->kp.omega_every_step = o.omega_every_step ? 1 : 0;
  s->kp.accepts = s->d_accepts;
  s->kp.overflow = s->d_overflow;
  s->kp.ctr = s->d_ctr;
  if (hipMemsetAsync(s->d_ctr, 0, 2 * sizeof(uint32_t), s->stream) != hipSuccess ||
      sync_kp(s) != SCMF_OK)
    return bail(fail(SCMF_E_HIP, "parameter upload failed"));
  hipLaunchKernelGGL(k_iota, dim3(grid_strided(s->n_chains)), dim3(BLOCK), 0, s->stream,
                     s->slot_chain, s->n_chains);
  hipLaunchKernelGGL(k_iota, dim3(grid_strided(s->n_chains)), dim3(BLOCK), 0, s->stream,
                     s->chain_slot, s->n_chains);
  hipLaunchKernelGGL(k_fill_btype, dim3(grid_strided(B)), dim3(BLOCK), 0, s->stream, s->kp, s->btype);
  choose_tiles(s);
  // Automatic ordering only where global atomics would cost more than the
  // reordering (below ~8M beads the atomic gridding takes < 0.3 ms).
  if (o.sort_every == 0 && s->n_beads < ((int64_t)1 << 23)) s->tile_lds = 0;
  // automatic period 100: chains drift slowly against the tiles (at 268M
  // beads the beads counted outside their windows stay at ~0.31% whether
  // the chains are reordered every 12, 25, 50 or 100 sweeps), so the
  // ordering's cost decides (5.88, 5.83, 5.81 ms a step at 25, 50, 100)
  s->sort_every = o.sort_every == 0 ? (s->tile_lds ? 100 : -1) : o.sort_every;
  if (hipMemsetAsync(s->d_fallback, 0, sizeof(unsigned long long), s->stream) != hipSuccess ||
      hipMemsetAsync(s->d_accepts, 0, sizeof(unsigned long long) * ACC_SLOTS, s->stream) != hipSuccess ||
      hipMemsetAsync(s->d_overflow, 0, sizeof(uint32_t), s->stream) != hipSuccess ||
      hipMemsetAsync(s->omega, 0, sizeof(float) * TC, s->stream) != hipSuccess ||
      hipMemsetAsync(s->counts, 0, sizeof(uint32_t) * s->count_words, s->stream) != hipSuccess)
    return bail(fail(SCMF_E_HIP, "initial clear failed"));
  if (o.init_coords) {
    hipLaunchKernelGGL(k_init_coords, dim3(grid_for(s->n_chains)), dim3(BLOCK), 0, s->stream,
                       (const KParams*)s->d_kp, (const BlockRef*)nullptr);
    if (hipGetLastError() != hipSuccess) return bail(fail(SCMF_E_HIP, "init launch failed"));
  } else {
    if (hipMemsetAsync(s->x, 0, sizeof(float) * B, s->stream) != hipSuccess ||
        hipMemsetAsync(s->y, 0, sizeof(float) * B, s->stream) != hipSuccess ||
        hipMemsetAsync(s->z, 0, sizeof(float) * B, s->stream) != hipSuccess)
      return bail(fail(SCMF_E_HIP, "position clear failed"));
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess)
    return bail(fail(SCMF_E_HIP, "initialisation failed on the device"));
  *out = s;
  return SCMF_OK;
}

}  // namespace

// ============================================================== C ABI
namespace {

// Checkpoint file: header, positions (fp32 [bead][xyz], reference order),
// counts (uint32 [type][cell]), omega (fp32 [type][cell]); little-endian.
struct CheckpointHeader {
  char magic[8];
  uint32_t version, T;
  int64_t n_beads, n_chains, n_cells;
  int32_t n[3], density_bits;
  uint64_t seed;
  uint32_t system_id, pad;
  uint64_t next_sweep;
  int64_t step;
  uint64_t desc_hash;
};
const char kCheckpointMagic[8] = {'S', 'C', 'M', 'F', 'C', 'K', 'P', 'T'};

// FNV-1a over everything that defines the system.
struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const uint8_t* b = (const uint8_t*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  }
  template <class T>
  void vec(const std::vector<T>& v) {
    const uint64_t n = v.size();
    add(&n, sizeof n);
    if (n) add(v.data(), sizeof(T) * v.size());
  }
};

uint64_t desc_hash(const HostDesc& d) {
  Fnv f;
  f.add(&d.T, sizeof d.T);
  f.vec(d.xn);
  f.add(d.L, sizeof d.L);
  f.add(d.n, sizeof d.n);
  f.add(&d.dtNxi, sizeof d.dtNxi);
  f.add(&d.nref, sizeof d.nref);
  f.vec(d.poly_length);
  f.vec(d.n_polymers);
  f.vec(d.monotype);
  f.vec(d.connection);
  return f.h;
}

}  // namespace

// ------------------------------------------------ asynchronous output
// Two snapshot slots.  A call takes a free slot (waiting for the oldest file
// if both are taken), copies counts and omega into the slot's device buffers
// on the system's stream and records an event; the writer thread waits for
// the event, copies the slot out to pinned host memory on its own stream and
// writes the file, then frees the slot.  The step loop only pays for the two
// device-to-device copies.
struct AsyncOut {
  static constexpr int SLOTS = 2;
  struct Slot {
    uint32_t* d_counts = nullptr;
    float* d_omega = nullptr;
    uint32_t* h_counts = nullptr;  // pinned
    float* h_omega = nullptr;      // pinned
    hipEvent_t ready = nullptr;
    bool busy = false;
  };
  struct Job {
    int slot;
    std::string path;  // density_<step>.dat
    std::string vtk;   // the VTK file of it, or empty
  };
  Slot slot[SLOTS];
  int device = 0;
  hipStream_t stream = nullptr;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> q;
  bool stop = false;
  std::string error;  // first failure since the last report
  // what a snapshot holds
  int64_t count_words = 0, TC = 0, n_cells = 0;
  int T = 0, bits = 32;
  int32_t n[3] = {0, 0, 0};
  double L[3] = {0, 0, 0};
  double fs = 1;

  std::string write_one(const Job& j) {
    Slot& sl = slot[j.slot];
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipEventSynchronize(sl.ready);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sl.h_counts, sl.d_counts, sizeof(uint32_t) * count_words,
                         hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sl.h_omega, sl.d_omega, sizeof(float) * TC, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return std::string("asynchronous output: ") + hipGetErrorString(e);
    const uint32_t* counts = sl.h_counts;
    std::vector<uint32_t> wide;
    if (bits == 16) {
      wide.resize((size_t)TC);
      for (int64_t i = 0; i < TC; ++i) wide[i] = (sl.h_counts[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
      counts = wide.data();
    }
    std::string err = write_density_file(j.path.c_str(), n, L, T, n_cells, counts, sl.h_omega, fs);
    if (err.empty() && !j.vtk.empty())
      err = write_vtk_file(j.vtk.c_str(), j.path.c_str(), n, L, T, n_cells, counts, sl.h_omega, fs);
    return err;
  }

  void run() {
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !q.empty(); });
        if (q.empty()) return;
        j = q.front();
        q.pop_front();
      }
      std::string e = write_one(j);
      {
        std::lock_guard<std::mutex> lk(mu);
        if (!e.empty() && error.empty()) error = e;
        slot[j.slot].busy = false;
      }
      cv.notify_all();
    }
  }

  bool idle() const {
    if (!q.empty()) return false;
    for (const Slot& sl : slot)
      if (sl.busy) return false;
    return true;
  }

  std::string take_error() {
    std::string e;
    e.swap(error);
    return e;
  }

  // Drain the queue and stop the thread; frees everything.
  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    if (th.joinable()) th.join();
    for (Slot& sl : slot) {
      if (sl.d_counts) (void)hipFree(sl.d_counts);
      if (sl.d_omega) (void)hipFree(sl.d_omega);
      if (sl.h_counts) (void)hipHostFree(sl.h_counts);
      if (sl.h_omega) (void)hipHostFree(sl.h_omega);
      if (sl.ready) (void)hipEventDestroy(sl.ready);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static int make_async_out(scmf_sys* s) {
  AsyncOut* o = new AsyncOut();
  o->device = s->device;
  o->count_words = s->count_words;
  o->TC = (int64_t)s->T * s->n_cells;
  o->n_cells = s->n_cells;
  o->T = s->T;
  o->bits = s->opts.density_bits;
  for (int a = 0; a < 3; ++a) {
    o->n[a] = s->desc.n[a];
    o->L[a] = s->desc.L[a];
  }
  o->fs = s->kp.fst[0];
  hipError_t e = hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking);
  for (AsyncOut::Slot& sl : o->slot) {
    if (e == hipSuccess) e = hipMalloc(&sl.d_counts, sizeof(uint32_t) * o->count_words);
    if (e == hipSuccess) e = hipMalloc(&sl.d_omega, sizeof(float) * o->TC);
    if (e == hipSuccess) e = hipHostMalloc(&sl.h_counts, sizeof(uint32_t) * o->count_words, 0);
    if (e == hipSuccess) e = hipHostMalloc(&sl.h_omega, sizeof(float) * o->TC, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    o->shutdown();
    delete o;
    return fail(SCMF_E_HIP, std::string("asynchronous output buffers: ") + hipGetErrorString(e));
  }
  o->th = std::thread([o] { o->run(); });
  s->out = o;
  return SCMF_OK;
}

extern "C" {

const char* scmf_last_error(void) { return g_err.c_str(); }
const char* scmf_version(void) { return "scmf-mi355x 0.1 (gfx950)"; }

int scmf_create(scmf_sys** out, const scmf_desc* desc, const scmf_opts* opts) {
  if (!desc) return fail(SCMF_E_ARG, "desc is null");
  std::string v = validate_desc(desc);
  if (!v.empty()) return fail(SCMF_E_ARG, v);
  HostDesc hd;
  hd.T = desc->n_mono_types;
  hd.xn.assign(desc->xn, desc->xn + (size_t)hd.T * hd.T);
  for (int i = 0; i < 3; ++i) {
    hd.L[i] = desc->L[i];
    hd.n[i] = desc->n[i];
  }
  hd.dtNxi = desc->dtNxi;
  hd.nref = desc->reference_Nbeads;
  hd.poly_length.assign(desc->poly_length, desc->poly_length + desc->n_poly_types);
  hd.n_polymers.assign(desc->n_polymers, desc->n_polymers + desc->n_poly_types);
  size_t msum = 0, csum = 0;
  for (int32_t N : hd.poly_length) {
    msum += N;
    csum += (size_t)N * N;
  }
  hd.monotype.assign(desc->monotype, desc->monotype + msum);
  hd.connection.assign(desc->connection, desc->connection + csum);
  return create_impl(out, hd, opts);
}

int scmf_create_from_files(scmf_sys** out, const char* configuration_dat, const char* polymer_dat,
                           const scmf_opts* opts) {
  if (!configuration_dat || !polymer_dat) return fail(SCMF_E_ARG, "file path is null");
  HostDesc hd;
  std::string e = read_polymer(polymer_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  e = read_configuration(configuration_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return create_impl(out, hd, opts);
}

void scmf_destroy(scmf_sys* s) {
  if (!s) return;
  if (s->out) {  // queued files are still written
    s->out->shutdown();
    delete s->out;
    s->out = nullptr;
  }
  (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
}

int scmf_mc_sweep(scmf_sys* s, int32_t n_sweeps) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (n_sweeps < 0) return fail(SCMF_E_ARG, "n_sweeps must be non-negative");
  return sweeps_impl(s, n_sweeps, false);
}

int scmf_grid_density(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  return grid_impl(s, true, false);
}

int scmf_sort_chains(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  return sort_chains(s);
}

int scmf_update_omega(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  Timed t;
  if (s->timing) {
    int rc = s->ev_begin(K_OMEGA, &t);
    if (rc) return rc;
  }
  if (s->n_cells % 4 == 0) {
    const void* fn = s->opts.density_bits == 32 ? omega4_kernel<32>(s->T) : omega4_kernel<16>(s->T);
    const KParams* kpp = s->d_kp;
    int cadence = 0;
    void* args[] = {(void*)&kpp, (void*)&cadence};
    HIP_TRY(hipLaunchKernel(fn, dim3(grid_strided(s->n_cells / 4)), dim3(BLOCK), args, 0,
                            s->stream));
  } else {
    hipLaunchKernelGGL(k_omega, dim3(grid_strided(s->n_cells)), dim3(BLOCK), 0, s->stream,
                       (const KParams*)s->d_kp, (const BlockRef*)nullptr, 0);
  }
  HIP_TRY(hipGetLastError());
  if (s->timing) return s->ev_end(&t);
  return SCMF_OK;
}

int scmf_step(scmf_sys* s, int32_t step) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (s->tile_lds && s->sort_every > 0 && s->sweeps % (uint64_t)s->sort_every == 0) {
    int rc = sort_chains(s);
    if (rc) return rc;
  }
  // the sweep clears the counts and the gridding advances the stream word
  // (two launches fewer than scmf_mc_sweep + scmf_grid_density)
  int rc = sweeps_impl(s, 1, true);
  if (rc) return rc;
  if ((rc = grid_impl(s, false, true))) return rc;
  if (s->opts.omega_every_step || (step % 10)) return scmf_update_omega(s);
  return SCMF_OK;
}

int scmf_copy_out(scmf_sys* s, int32_t what, void* host, size_t bytes) {
  if (!s || !host) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  if (what == SCMF_POS) {
    if (bytes < sizeof(float) * 3 * (size_t)s->n_beads) return fail(SCMF_E_ARG, "buffer too small");
    int rc = ensure_staging(s, std::min<int64_t>(s->n_beads, STAGE_BEADS));
    if (rc) return rc;
    for (int64_t b0 = 0; b0 < s->n_beads; b0 += STAGE_BEADS) {
      const int64_t n = std::min<int64_t>(STAGE_BEADS, s->n_beads - b0);
      hipLaunchKernelGGL(k_gather_ref, dim3(grid_strided(n)), dim3(BLOCK), 0, s->stream, s->kp,
                         s->x, s->y, s->z, s->staging, b0, n);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync((float*)host + 3 * b0, s->staging, sizeof(float) * 3 * n,
                             hipMemcpyDeviceToHost, s->stream));
      HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return SCMF_OK;
  }
  if (what == SCMF_DENSITY) {
    if (bytes < sizeof(uint32_t) * (size_t)TC) return fail(SCMF_E_ARG, "buffer too small");
    if (s->opts.density_bits == 32) {
      HIP_TRY(hipMemcpyAsync(host, s->counts, sizeof(uint32_t) * TC, hipMemcpyDeviceToHost, s->stream));
    } else {
      uint32_t* wide = nullptr;
      HIP_TRY(hipMalloc(&wide, sizeof(uint32_t) * TC));
      hipLaunchKernelGGL(k_widen<16>, dim3(grid_strided(TC)), dim3(BLOCK), 0, s->stream, s->counts,
                         wide, TC);
      hipError_t e1 = hipMemcpyAsync(host, wide, sizeof(uint32_t) * TC, hipMemcpyDeviceToHost, s->stream);
      hipError_t e2 = hipStreamSynchronize(s->stream);
      (void)hipFree(wide);
      HIP_TRY(e1);
      HIP_TRY(e2);
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SCMF_OK;
  }
  if (what == SCMF_OMEGA) {
    if (bytes < sizeof(float) * (size_t)TC) return fail(SCMF_E_ARG, "buffer too small");
    HIP_TRY(hipMemcpyAsync(host, s->omega, sizeof(float) * TC, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SCMF_OK;
  }
  return fail(SCMF_E_ARG, "unknown copy_out selector");
}

}  // extern "C"

namespace {

// The device address of `scratch` on each GPU, looked up once per device with
// that device current (the array's name is a host-side stand-in, not a
// device address), and the lock that serialises the calls using it.
constexpr int NOALLOC_MAX_DEVICES = 64;
uint32_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int scratch_on(int device, uint32_t** out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(SCMF_E_NODEV, "device index out of range");
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(SCMF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (uint32_t*)p;
  }
  *out = g_scratch_ptr[device];
  return SCMF_OK;
}

}  // namespace

extern "C" {

int scmf_copy_out_noalloc(scmf_sys* s, int32_t what, void* host, size_t bytes) {
  if (!s || !host) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (what == SCMF_POS) {
    if (bytes < sizeof(float) * 3 * (size_t)s->n_beads) return fail(SCMF_E_ARG, "buffer too small");
    if (3 * s->n_beads > NOALLOC_WORDS)
      return fail(SCMF_E_ARG, "scmf_copy_out_noalloc: positions of at most " +
                                  std::to_string(NOALLOC_WORDS / 3) + " beads (the system has " +
                                  std::to_string(s->n_beads) + "); use scmf_copy_out");
    uint32_t* ptr = nullptr;
    int rc = scratch_on(s->device, &ptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather_ref, dim3(grid_strided(s->n_beads)), dim3(BLOCK), 0, s->stream,
                       s->kp, s->x, s->y, s->z, (float*)ptr, (int64_t)0, s->n_beads);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host, ptr, sizeof(float) * 3 * s->n_beads, hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SCMF_OK;
  }
  if (what == SCMF_DENSITY) {
    if (bytes < sizeof(uint32_t) * (size_t)TC) return fail(SCMF_E_ARG, "buffer too small");
    if (s->opts.density_bits == 32) {
      HIP_TRY(hipMemcpyAsync(host, s->counts, sizeof(uint32_t) * TC, hipMemcpyDeviceToHost, s->stream));
    } else {
      if (TC > NOALLOC_WORDS)
        return fail(SCMF_E_ARG, "scmf_copy_out_noalloc: uint16 counts of at most " +
                                    std::to_string(NOALLOC_WORDS) + " (types x cells; the system has " +
                                    std::to_string(TC) + "); use scmf_copy_out");
      uint32_t* ptr = nullptr;
      int rc = scratch_on(s->device, &ptr);
      if (rc) return rc;
      hipLaunchKernelGGL(k_widen<16>, dim3(grid_strided(TC)), dim3(BLOCK), 0, s->stream, s->counts,
                         ptr, TC);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(host, ptr, sizeof(uint32_t) * TC, hipMemcpyDeviceToHost, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SCMF_OK;
  }
  if (what == SCMF_OMEGA) {
    if (bytes < sizeof(float) * (size_t)TC) return fail(SCMF_E_ARG, "buffer too small");
    HIP_TRY(hipMemcpyAsync(host, s->omega, sizeof(float) * TC, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SCMF_OK;
  }
  return fail(SCMF_E_ARG, "unknown copy_out selector");
}

int scmf_copy_in_pos(scmf_sys* s, const float* xyz, size_t n_beads) {
  if (!s || !xyz) return fail(SCMF_E_ARG, "null argument");
  s->cells_fresh = false;
  if ((int64_t)n_beads != s->n_beads) return fail(SCMF_E_ARG, "bead count mismatch");
  int rc = ensure_staging(s, std::min<int64_t>(s->n_beads, STAGE_BEADS));
  if (rc) return rc;
  for (int64_t b0 = 0; b0 < s->n_beads; b0 += STAGE_BEADS) {
    const int64_t n = std::min<int64_t>(STAGE_BEADS, s->n_beads - b0);
    HIP_TRY(hipMemcpyAsync(s->staging, xyz + 3 * b0, sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                           s->stream));
    hipLaunchKernelGGL(k_scatter_ref, dim3(grid_strided(n)), dim3(BLOCK), 0, s->stream, s->kp,
                       s->staging, s->x, s->y, s->z, b0, n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  return SCMF_OK;
}

int scmf_copy_in_omega(scmf_sys* s, const float* omega, size_t n_values) {
  if (!s || !omega) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  if ((int64_t)n_values != TC) return fail(SCMF_E_ARG, "omega size mismatch");
  HIP_TRY(hipMemcpyAsync(s->omega, omega, sizeof(float) * TC, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

int scmf_copy_in_density(scmf_sys* s, const uint32_t* counts, size_t n_values) {
  if (!s || !counts) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  if ((int64_t)n_values != TC) return fail(SCMF_E_ARG, "density size mismatch");
  if (s->opts.density_bits == 32) {
    HIP_TRY(hipMemcpyAsync(s->counts, counts, sizeof(uint32_t) * TC, hipMemcpyHostToDevice, s->stream));
  } else {
    std::vector<uint32_t> packed(s->count_words, 0u);
    for (int64_t i = 0; i < TC; ++i) {
      if (counts[i] > 0xFFFFu) return fail(SCMF_E_OVERFLOW, "count above 65535 for uint16 fields");
      packed[i >> 1] |= counts[i] << ((i & 1) * 16);
    }
    HIP_TRY(hipMemcpyAsync(s->counts, packed.data(), sizeof(uint32_t) * s->count_words,
                           hipMemcpyHostToDevice, s->stream));
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

int scmf_set_rng(scmf_sys* s, uint64_t seed, uint64_t next_sweep) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (next_sweep >= 0xFFFFFFFFull) return fail(SCMF_E_ARG, "sweep counter exhausted");
  s->opts.seed = seed;
  s->kp.seed = seed;
  s->sweeps = next_sweep;
  int rc = sync_kp(s);
  if (rc) return rc;
  return set_ctr(s, (uint32_t)next_sweep);
}

int scmf_structure_factor(scmf_sys* s, const double* coef, double dq, int32_t by_q2,
                          int32_t n_bins, double* s_sum, int64_t* n_modes) {
  if (!s || !coef || !s_sum || !n_modes) return fail(SCMF_E_ARG, "null argument");
  if (!(dq > 0) || !std::isfinite(dq)) return fail(SCMF_E_ARG, "dq must be positive and finite");
  if (n_bins < 1 || n_bins > SCMF_SQ_MAX_BINS) return fail(SCMF_E_ARG, "n_bins out of range");
  if (by_q2 != 0 && by_q2 != 1) return fail(SCMF_E_ARG, "by_q2 must be 0 or 1");
  SqCoef cf{};
  for (int t = 0; t < s->T; ++t) {
    if (!std::isfinite(coef[t])) return fail(SCMF_E_ARG, "coefficients must be finite");
    cf.c[t] = coef[t];
  }
  const int nx = s->kp.nx, ny = s->kp.ny, nz = s->kp.nz;
  const int64_t hx = nx / 2 + 1;
  if (!s->fft_ready) {
    const size_t bytes = (size_t)nz * ny * hx * 2 * sizeof(double);
    if (!s->fft_buf) HIP_TRY(hipMalloc(&s->fft_buf, bytes));
    hipfftHandle h = 0;
    if (hipfftCreate(&h) != HIPFFT_SUCCESS) return fail(SCMF_E_HIP, "hipfftCreate failed");
    long long n[3] = {nz, ny, nx};
    long long inembed[3] = {nz, ny, 2 * hx};
    long long onembed[3] = {nz, ny, hx};
    size_t work = 0;
    hipfftResult r = hipfftMakePlanMany64(h, 3, n, inembed, 1, (long long)nz * ny * 2 * hx,
                                          onembed, 1, (long long)nz * ny * hx, HIPFFT_D2Z, 1,
                                          &work);
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(h, s->stream);
    if (r != HIPFFT_SUCCESS) {
      (void)hipfftDestroy(h);
      return fail(SCMF_E_HIP, "hipfft plan for the structure factor failed (" +
                                  std::to_string((int)r) + ")");
    }
    s->fft = h;
    s->fft_ready = true;
  }
  if (s->sq_cap < n_bins) {
    if (s->d_sq_sum) HIP_TRY(hipFree(s->d_sq_sum));
    if (s->d_sq_cnt) HIP_TRY(hipFree(s->d_sq_cnt));
    s->d_sq_sum = nullptr;
    s->d_sq_cnt = nullptr;
    s->sq_cap = 0;
    HIP_TRY(hipMalloc(&s->d_sq_sum, (size_t)n_bins * sizeof(double)));
    HIP_TRY(hipMalloc(&s->d_sq_cnt, (size_t)n_bins * sizeof(unsigned long long)));
    s->sq_cap = n_bins;
  }
  hipLaunchKernelGGL(k_sq_weights, dim3(grid_strided(s->n_cells)), dim3(BLOCK), 0, s->stream,
                     s->kp, cf, s->fft_buf, 2 * hx);
  HIP_TRY(hipGetLastError());
  const hipfftResult r =
      hipfftExecD2Z(s->fft, (hipfftDoubleReal*)s->fft_buf, (hipfftDoubleComplex*)s->fft_buf);
  if (r != HIPFFT_SUCCESS)
    return fail(SCMF_E_HIP, "hipfftExecD2Z failed (" + std::to_string((int)r) + ")");
  HIP_TRY(hipMemsetAsync(s->d_sq_sum, 0, (size_t)n_bins * sizeof(double), s->stream));
  HIP_TRY(hipMemsetAsync(s->d_sq_cnt, 0, (size_t)n_bins * sizeof(unsigned long long), s->stream));
  const double two_pi = 6.283185307179586;
  const int64_t modes = (int64_t)nz * ny * hx;
  int64_t blocks = (modes + BLOCK - 1) / BLOCK;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_sq_shells, dim3((unsigned)blocks), dim3(BLOCK), 0, s->stream,
                     (const double*)s->fft_buf, nx, ny, nz, two_pi / s->kp.Lx, two_pi / s->kp.Ly,
                     two_pi / s->kp.Lz, 1.0 / dq, by_q2, 1.0 / (double)s->n_beads, n_bins,
                     s->d_sq_sum, s->d_sq_cnt);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(s_sum, s->d_sq_sum, (size_t)n_bins * sizeof(double),
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(n_modes, s->d_sq_cnt, (size_t)n_bins * sizeof(int64_t),
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

int scmf_stats(scmf_sys* s, scmf_stats_t* out) {
  if (!s || !out) return fail(SCMF_E_ARG, "null argument");
  const unsigned nb = grid_for(s->n_chains);
  hipLaunchKernelGGL(k_re2, dim3(nb), dim3(BLOCK), 0, s->stream, s->kp, s->x, s->y, s->z,
                     s->d_partial);
  HIP_TRY(hipGetLastError());
  std::vector<double> part(nb);
  unsigned long long fb = 0;
  std::vector<unsigned long long> accs(ACC_SLOTS);
  uint32_t ovf = 0;
  HIP_TRY(hipMemcpyAsync(part.data(), s->d_partial, sizeof(double) * nb, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(accs.data(), s->d_accepts, sizeof(unsigned long long) * ACC_SLOTS,
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(&ovf, s->d_overflow, sizeof ovf, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(&fb, s->d_fallback, sizeof fb, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  double sum = 0;
  for (double v : part) sum += v;
  unsigned long long acc = 0;
  for (unsigned long long v : accs) acc += v;
  out->moves = s->moves;
  out->accepts = acc;
  out->sweeps = s->sweeps;
  out->re2 = sum / (double)s->n_chains;
  out->overflow = ovf ? 1 : 0;
  out->grid_fallback = fb;
  out->sorts = s->n_sorts;
  out->cell_grids = s->n_cell_grids;
  return SCMF_OK;
}

int scmf_reset_stats(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  HIP_TRY(hipMemsetAsync(s->d_accepts, 0, sizeof(unsigned long long) * ACC_SLOTS, s->stream));
  HIP_TRY(hipMemsetAsync(s->d_overflow, 0, sizeof(uint32_t), s->stream));
  HIP_TRY(hipMemsetAsync(s->d_fallback, 0, sizeof(unsigned long long), s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->moves = 0;
  return SCMF_OK;
}

int scmf_enable_timing(scmf_sys* s, int32_t on) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  s->timing = on != 0;
  return SCMF_OK;
}

int scmf_timing(scmf_sys* s, scmf_timing_t* out) {
  if (!s || !out) return fail(SCMF_E_ARG, "null argument");
  int rc = s->drain_timing();
  if (rc) return rc;
  out->sweep_n = s->t_n[K_SWEEP];
  out->grid_n = s->t_n[K_GRID];
  out->omega_n = s->t_n[K_OMEGA];
  out->sweep_ms = s->t_n[K_SWEEP] ? s->t_sum[K_SWEEP] / s->t_n[K_SWEEP] : 0;
  out->grid_ms = s->t_n[K_GRID] ? s->t_sum[K_GRID] / s->t_n[K_GRID] : 0;
  out->omega_ms = s->t_n[K_OMEGA] ? s->t_sum[K_OMEGA] / s->t_n[K_OMEGA] : 0;
  return SCMF_OK;
}

int scmf_reset_timing(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  int rc = s->drain_timing();
  if (rc) return rc;
  for (int k = 0; k < K_KINDS; ++k) {
    s->t_sum[k] = 0;
    s->t_n[k] = 0;
  }
  return SCMF_OK;
}

int scmf_autotune(scmf_sys* s, int32_t sweeps_per_variant, int32_t* chosen, float* best_ms) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (sweeps_per_variant <= 0) sweeps_per_variant = 2;
  const int64_t B = s->n_beads;
  float* backup = nullptr;
  if (hipMalloc(&backup, sizeof(float) * 3 * B) != hipSuccess) {
    (void)hipGetLastError();
    return fail(SCMF_E_NOMEM, "autotune needs a copy of the positions");
  }
  std::vector<unsigned long long> acc0(ACC_SLOTS);
  HIP_TRY(hipMemcpyAsync(acc0.data(), s->d_accepts, sizeof(unsigned long long) * ACC_SLOTS,
                         hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipMemcpyAsync(backup, s->x, sizeof(float) * B, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(backup + B, s->y, sizeof(float) * B, hipMemcpyDeviceToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(backup + 2 * B, s->z, sizeof(float) * B, hipMemcpyDeviceToDevice, s->stream));
  const uint64_t sweeps0 = s->sweeps, moves0 = s->moves;
  const SweepVariant* keep = s->variant;
  const size_t keep_lds = s->lds_bytes;
  const int32_t keep_req = s->opts.lanes_per_chain;
  const SweepVariant* best = keep;
  size_t best_lds = keep_lds;
  float best_t = 1e30f;
  hipEvent_t a, b;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  int rc = SCMF_OK;
  for (int32_t L : {-1, 1, 2, 4, 8, 16}) {
    s->opts.lanes_per_chain = L;
    if (choose_sweep(s) != SCMF_OK) continue;
    if ((rc = launch_sweep(s))) break;  // warm-up launch
    HIP_TRY(hipEventRecord(a, s->stream));
    for (int32_t k = 0; k < sweeps_per_variant && !rc; ++k) rc = launch_sweep(s);
    HIP_TRY(hipEventRecord(b, s->stream));
    HIP_TRY(hipEventSynchronize(b));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, a, b));
    ms /= (float)sweeps_per_variant;
    if (ms < best_t) {
      best_t = ms;
      best = s->variant;
      best_lds = s->lds_bytes;
    }
    HIP_TRY(hipMemcpyAsync(s->x, backup, sizeof(float) * B, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->y, backup + B, sizeof(float) * B, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->z, backup + 2 * B, sizeof(float) * B, hipMemcpyDeviceToDevice, s->stream));
    if ((rc = set_ctr(s, (uint32_t)sweeps0))) break;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIP_TRY(hipMemcpyAsync(s->d_accepts, acc0.data(), sizeof(unsigned long long) * ACC_SLOTS,
                         hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  (void)hipFree(backup);
  s->cells_fresh = false;  // positions restored
  s->sweeps = sweeps0;
  s->moves = moves0;
  s->opts.lanes_per_chain = keep_req;
  if (rc) {
    s->variant = keep;
    s->lds_bytes = keep_lds;
    return rc;
  }
  s->variant = best;
  s->lds_bytes = best_lds;
  if (chosen) *chosen = best->L == 0 ? -1 : best->L;
  if (best_ms) *best_ms = best_t;
  return SCMF_OK;
}

int scmf_sync(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return SCMF_OK;
}

int scmf_sizes(scmf_sys* s, int64_t* n_beads, int64_t* n_chains, int64_t* n_cells, int32_t* n_types) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (n_beads) *n_beads = s->n_beads;
  if (n_chains) *n_chains = s->n_chains;
  if (n_cells) *n_cells = s->n_cells;
  if (n_types) *n_types = s->T;
  return SCMF_OK;
}

int scmf_write_density_async(scmf_sys* s, const char* dir, int32_t step) {
  return scmf_write_outputs_async(s, dir, step, nullptr);
}

int scmf_write_outputs_async(scmf_sys* s, const char* dir, int32_t step, const char* vtk_path) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (!s->out) {
    int rc = make_async_out(s);
    if (rc) return rc;
  }
  AsyncOut* o = s->out;
  std::string path = density_file_name(step);
  if (dir && *dir) path = std::string(dir) + "/" + path;
  int k = -1;
  {
    std::unique_lock<std::mutex> lk(o->mu);
    o->cv.wait(lk, [&] {
      for (const AsyncOut::Slot& sl : o->slot)
        if (!sl.busy) return true;
      return false;
    });
    const std::string e = o->take_error();
    if (!e.empty()) return fail(SCMF_E_IO, e);
    for (int i = 0; i < AsyncOut::SLOTS; ++i)
      if (!o->slot[i].busy) {
        k = i;
        break;
      }
    o->slot[k].busy = true;
  }
  AsyncOut::Slot& sl = o->slot[k];
  hipError_t e = hipMemcpyAsync(sl.d_counts, s->counts, sizeof(uint32_t) * s->count_words,
                                hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(sl.d_omega, s->omega, sizeof(float) * o->TC, hipMemcpyDeviceToDevice,
                       s->stream);
  if (e == hipSuccess) e = hipEventRecord(sl.ready, s->stream);
  {
    std::lock_guard<std::mutex> lk(o->mu);
    if (e == hipSuccess)
      o->q.push_back({k, path, vtk_path ? std::string(vtk_path) : std::string()});
    else
      sl.busy = false;
  }
  o->cv.notify_all();
  HIP_TRY(e);
  return SCMF_OK;
}

int scmf_output_wait(scmf_sys* s) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  if (!s->out) return SCMF_OK;
  AsyncOut* o = s->out;
  std::unique_lock<std::mutex> lk(o->mu);
  o->cv.wait(lk, [&] { return o->idle(); });
  const std::string e = o->take_error();
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

int scmf_write_density(scmf_sys* s, const char* dir, int32_t step) {
  if (!s) return fail(SCMF_E_ARG, "null handle");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  std::vector<uint32_t> counts(TC);
  std::vector<float> om(TC);
  int rc = scmf_copy_out(s, SCMF_DENSITY, counts.data(), sizeof(uint32_t) * TC);
  if (rc) return rc;
  if ((rc = scmf_copy_out(s, SCMF_OMEGA, om.data(), sizeof(float) * TC))) return rc;
  std::string path = density_file_name(step);
  if (dir && *dir) path = std::string(dir) + "/" + path;
  std::string e = write_density_file(path.c_str(), s->desc.n, s->desc.L, s->T, s->n_cells,
                                     counts.data(), om.data(), s->kp.fst[0]);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

int scmf_write_coord(scmf_sys* s, const char* path) {
  if (!s || !path) return fail(SCMF_E_ARG, "null argument");
  std::vector<float> xyz(3 * (size_t)s->n_beads);
  int rc = scmf_copy_out(s, SCMF_POS, xyz.data(), sizeof(float) * xyz.size());
  if (rc) return rc;
  std::string e = write_coord_file(path, xyz.data(), s->n_beads);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

int scmf_checkpoint_write(scmf_sys* s, const char* path, int64_t step) {
  if (!s || !path) return fail(SCMF_E_ARG, "null argument");
  const int64_t TC = (int64_t)s->T * s->n_cells;
  CheckpointHeader h{};
  memcpy(h.magic, kCheckpointMagic, 8);
  h.version = 1;
  h.T = (uint32_t)s->T;
  h.n_beads = s->n_beads;
  h.n_chains = s->n_chains;
  h.n_cells = s->n_cells;
  for (int a = 0; a < 3; ++a) h.n[a] = s->desc.n[a];
  h.density_bits = s->opts.density_bits;
  h.seed = s->opts.seed;
  h.system_id = s->opts.system_id;
  h.next_sweep = s->sweeps;
  h.step = step;
  h.desc_hash = desc_hash(s->desc);
  std::vector<float> pos(3 * (size_t)s->n_beads);
  std::vector<uint32_t> counts((size_t)TC);
  std::vector<float> om((size_t)TC);
  int rc;
  if ((rc = scmf_copy_out(s, SCMF_POS, pos.data(), sizeof(float) * pos.size())) ||
      (rc = scmf_copy_out(s, SCMF_DENSITY, counts.data(), sizeof(uint32_t) * counts.size())) ||
      (rc = scmf_copy_out(s, SCMF_OMEGA, om.data(), sizeof(float) * om.size())))
    return rc;
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return fail(SCMF_E_IO, "cannot open " + tmp + " for writing");
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(pos.data(), sizeof(float), pos.size(), f) == pos.size() &&
            fwrite(counts.data(), sizeof(uint32_t), counts.size(), f) == counts.size() &&
            fwrite(om.data(), sizeof(float), om.size(), f) == om.size();
  ok = (fclose(f) == 0) && ok;
  if (!ok || rename(tmp.c_str(), path) != 0) {
    remove(tmp.c_str());
    return fail(SCMF_E_IO, std::string("writing the checkpoint ") + path + " failed");
  }
  return SCMF_OK;
}

int scmf_checkpoint_read(scmf_sys* s, const char* path, int64_t* step) {
  if (!s || !path) return fail(SCMF_E_ARG, "null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(SCMF_E_IO, std::string("cannot open ") + path);
  CheckpointHeader h{};
  const int64_t TC = (int64_t)s->T * s->n_cells;
  auto bad = [&](const std::string& why) {
    fclose(f);
    return fail(SCMF_E_IO, std::string(path) + ": " + why);
  };
  if (fread(&h, sizeof h, 1, f) != 1) return bad("truncated header");
  if (memcmp(h.magic, kCheckpointMagic, 8) != 0 || h.version != 1) return bad("not a checkpoint file");
  if ((int)h.T != s->T || h.n_beads != s->n_beads || h.n_chains != s->n_chains ||
      h.n_cells != s->n_cells || h.desc_hash != desc_hash(s->desc))
    return bad("written for a different system description");
  if (h.density_bits != s->opts.density_bits) return bad("written with other density counters");
  if (h.next_sweep >= 0xFFFFFFFFull) return bad("sweep counter out of range");
  std::vector<float> pos(3 * (size_t)s->n_beads);
  std::vector<uint32_t> counts((size_t)TC);
  std::vector<float> om((size_t)TC);
  if (fread(pos.data(), sizeof(float), pos.size(), f) != pos.size() ||
      fread(counts.data(), sizeof(uint32_t), counts.size(), f) != counts.size() ||
      fread(om.data(), sizeof(float), om.size(), f) != om.size())
    return bad("truncated data");
  fclose(f);
  int rc;
  if ((rc = scmf_copy_in_pos(s, pos.data(), (size_t)s->n_beads)) ||
      (rc = scmf_copy_in_density(s, counts.data(), (size_t)TC)) ||
      (rc = scmf_copy_in_omega(s, om.data(), (size_t)TC)))
    return rc;
  s->opts.system_id = h.system_id;
  s->kp.system_id = h.system_id;
  if ((rc = scmf_set_rng(s, h.seed, h.next_sweep))) return rc;
  if (step) *step = h.step;
  return SCMF_OK;
}

namespace {
void fill_summary(const scmf_desc* d, scmf_summary* out) {
  memset(out, 0, sizeof *out);
  out->n_mono_types = d->n_mono_types;
  out->n_poly_types = d->n_poly_types;
  out->reference_Nbeads = d->reference_Nbeads;
  for (int i = 0; i < 3; ++i) {
    out->n[i] = d->n[i];
    out->L[i] = d->L[i];
  }
  out->dtNxi = d->dtNxi;
  int64_t chains = 0, beads = 0;
  for (int p = 0; p < d->n_poly_types; ++p) {
    chains += d->n_polymers[p];
    beads += d->n_polymers[p] * (int64_t)d->poly_length[p];
  }
  out->n_chains = chains;
  out->n_beads = beads;
  out->n_cells = (int64_t)d->n[0] * d->n[1] * d->n[2];
  const int tt = d->n_mono_types * d->n_mono_types;
  for (int i = 0; i < tt && i < 64; ++i) out->xn[i] = d->xn[i];
  const uint64_t TC = (uint64_t)d->n_mono_types * out->n_cells;
  int nmax = 0;
  for (int p = 0; p < d->n_poly_types; ++p) nmax = std::max(nmax, (int)d->poly_length[p]);
  const uint64_t B4 = (uint64_t)padded(beads);
  // positions, monomer types, slot maps, omega, partial sums and parameters
  uint64_t fixed = 12ull * B4 + (uint64_t)beads + 8ull * chains + 4ull * TC +
                   8ull * ((chains + BLOCK - 1) / BLOCK) + 4096;
  if (nmax <= 255 && chains >= 65536)  // schedule records
    fixed += (uint64_t)schedule_record_bytes(nmax, chains);
  if (beads >= ((int64_t)1 << 23)) {  // automatic ordering: second position buffer, two cell
    fixed += 20ull * B4 + 16ull * chains;  // buffers, keys and maps
    // window images and tile counters of the tiled gridding, from the same plan
    std::vector<std::pair<int, int64_t>> lens;
    for (int p = 0; p < d->n_poly_types; ++p)
      lens.push_back({(int)d->poly_length[p], d->n_polymers[p] * (int64_t)d->poly_length[p]});
    const int n[3] = {d->n[0], d->n[1], d->n[2]};
    const double dd[3] = {d->L[0] / d->n[0], d->L[1] / d->n[1], d->L[2] / d->n[2]};
    TileGeo g{};
    if (plan_tiles(n, dd, lens, d->reference_Nbeads, d->n_mono_types, &g))
      fixed += 2ull * ((uint64_t)g.ntiles * (uint64_t)g.slot_elems + 64) +
               12ull * (uint64_t)d->n_poly_types * (uint64_t)(g.ntiles + 1);
  }
  fixed += 27ull << 21;  // up to 27 device allocations, each rounded to 2 MiB
  out->device_bytes_16 = fixed + 4ull * ((TC + 1) / 2);
  out->device_bytes_32 = fixed + 4ull * TC;
}
}  // namespace

int scmf_inspect_desc(const scmf_desc* desc, scmf_summary* out) {
  if (!desc || !out) return fail(SCMF_E_ARG, "null argument");
  std::string v = validate_desc(desc);
  if (!v.empty()) return fail(SCMF_E_ARG, v);
  fill_summary(desc, out);
  return SCMF_OK;
}

int scmf_inspect_files(const char* configuration_dat, const char* polymer_dat, scmf_summary* out) {
  if (!configuration_dat || !polymer_dat || !out) return fail(SCMF_E_ARG, "null argument");
  HostDesc hd;
  std::string e = read_polymer(polymer_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  e = read_configuration(configuration_dat, &hd);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  scmf_desc d = hd.view();
  return scmf_inspect_desc(&d, out);
}

int scmf_write_density_host(const char* path, const int32_t n[3], const double L[3], int32_t T,
                            const uint32_t* counts, const float* omega, double field_scaling) {
  if (!path || !n || !L || !counts || !omega || T <= 0) return fail(SCMF_E_ARG, "null argument");
  const int64_t C = (int64_t)n[0] * n[1] * n[2];
  std::string e = write_density_file(path, n, L, T, C, counts, omega, field_scaling);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

int scmf_write_vtk_host(const char* vtk_path, const char* density_name, const int32_t n[3],
                        const double L[3], int32_t T, const uint32_t* counts, const float* omega,
                        double field_scaling) {
  if (!vtk_path || !density_name || !n || !L || !counts || !omega || T <= 0)
    return fail(SCMF_E_ARG, "null argument");
  const int64_t C = (int64_t)n[0] * n[1] * n[2];
  std::string e = write_vtk_file(vtk_path, density_name, n, L, T, C, counts, omega, field_scaling);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

int scmf_density_to_vtk(const char* density_path, const char* vtk_path) {
  if (!density_path || !vtk_path) return fail(SCMF_E_ARG, "null path");
  std::string e = density_to_vtk(density_path, vtk_path);
  if (!e.empty()) return fail(SCMF_E_IO, e);
  return SCMF_OK;
}

}  // extern "C"

// ============================================================== batches
// A batch advances many systems together: every phase of a step is one launch
// (per sweep-kernel shape) over the workgroups of all members, each workgroup
// looking up its member in a block table; the whole step is captured once in
// a HIP graph and replayed.  Members keep their own memory and can still be
// read or written through their handles between batch calls (after
// scmf_batch_sync).  A member may carry a stop: from that sweep count on its
// workgroups return at once, so members of different lengths share every
// step.
//
// The step graph is static across changes of membership and stops.  Its
// launches are sized to capacities (members, workgroups per phase, LDS rows
// per sweep shape) and read everything else from a device-side slot table:
// the members' parameter records and per-phase block tables.  Entries beyond
// the current membership name a vacant record (stop 0, sweep words zero), so
// their workgroups return at once.  A new membership that fits is installed
// by uploading the table on the batch stream, in order with the queued steps,
// without waiting for them and without a capture.  One that does not fit (a
// new sweep shape, longer chains, more members or workgroups, or far fewer,
// or a member gridded in tiles, whose launches name it) builds the next
// generation: its table in a stream-ordered allocation, its graph captured on
// a stream of its own with half again the room needed, the previous
// generation's memory freed in stream order and its graph destroyed once an
// event after its last launch has passed.
struct scmf_batch {
  std::vector<scmf_sys*> m;
  std::vector<uint32_t> stop;    // per member: the sweep count it stops at (NO_STOP: none)
  hipStream_t stream = nullptr;  // the steps
  hipStream_t cap = nullptr;     // graph capture only
  struct Group {
    const SweepVariant* v;
    int nmax;                    // LDS rows the launch is sized for
    size_t lds;
    size_t bt;                   // byte offset of its block table in d_buf
    unsigned nblocks;            // workgroups launched (the table's capacity)
  };
  // the current generation: the captured graph and the slot table it reads
  uint8_t* d_buf = nullptr;      // KParams of member_cap members and the vacant record,
                                 // the block tables, the vacant record's sweep words
  size_t buf_bytes = 0;
  int member_cap = 0;            // the vacant record is Ps[member_cap]
  bool tiled = false;            // the graph holds per-member tiled gridding launches
  std::vector<Group> groups;
  size_t bt_zero = 0, bt_grid = 0, bt_omega = 0, vac_ctr = 0;
  unsigned n_zero = 0, n_grid = 0, n_omega = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t uploads = 0;          // memberships installed into the current graph's table
  uint64_t used_blocks = 0;      // workgroups of the current members (of those launched)
  struct Retired {
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipEvent_t done;
  };
  std::vector<Retired> retired;
  struct Stage {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t used = nullptr;
    bool pending = false;
  };
  Stage stage[3];                // pinned staging of the table uploads, used in turn
  int next_stage = 0;
  int forced_L = 0;              // 0: each member's own shape
  uint64_t steps = 0;
  uint64_t generations = 0;
  // tickets: events around enqueued steps, timed from an origin event
  hipEvent_t origin = nullptr;
  std::chrono::steady_clock::time_point origin_host;
  struct Ticket {
    uint64_t id;
    hipEvent_t a, b;
  };
  std::deque<Ticket> tickets;
  uint64_t next_ticket = 1;
};

namespace {

// Destroy the retired graphs whose last launch has run (all of them, waiting,
// with `wait`).
void batch_reclaim(scmf_batch* b, bool wait) {
  std::vector<scmf_batch::Retired> keep;
  for (auto& r : b->retired) {
    const hipError_t q = wait ? hipEventSynchronize(r.done) : hipEventQuery(r.done);
    if (q == hipSuccess) {
      if (r.exec) (void)hipGraphExecDestroy(r.exec);
      if (r.graph) (void)hipGraphDestroy(r.graph);
      (void)hipEventDestroy(r.done);
    } else {
      (void)hipGetLastError();
      keep.push_back(r);
    }
  }
  b->retired.swap(keep);
}

// Retire the current generation: its memory is freed in stream order after the
// steps queued so far, its graph destroyed once they have run.
int batch_retire(scmf_batch* b) {
  if (b->exec || b->graph) {
    hipEvent_t ev = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(ev, b->stream));
    b->retired.push_back({b->graph, b->exec, ev});
    b->graph = nullptr;
    b->exec = nullptr;
  }
  if (b->d_buf) {
    HIP_TRY(hipFreeAsync(b->d_buf, b->stream));
    b->d_buf = nullptr;
  }
  b->groups.clear();
  b->buf_bytes = 0;
  b->member_cap = 0;
  b->tiled = false;
  b->n_zero = b->n_grid = b->n_omega = 0;
  return SCMF_OK;
}

// Copy host bytes to device memory on the batch stream through a pinned
// staging slot, so the host does not wait for the steps queued before it.
int batch_upload(scmf_batch* b, void* dst, const void* src, size_t bytes) {
  scmf_batch::Stage& st = b->stage[b->next_stage];
  b->next_stage = (b->next_stage + 1) % 3;
  if (st.pending) {
    HIP_TRY(hipEventSynchronize(st.used));
    st.pending = false;
  }
  if (st.cap < bytes) {
    const size_t cap = std::max(bytes, 2 * st.cap);
    if (st.host) (void)hipHostFree(st.host);
    st.host = nullptr;
    st.cap = 0;
    HIP_TRY(hipHostMalloc(&st.host, cap, hipHostMallocDefault));
    st.cap = cap;
  }
  if (!st.used) HIP_TRY(hipEventCreateWithFlags(&st.used, hipEventDisableTiming));
  memcpy(st.host, src, bytes);
  HIP_TRY(hipMemcpyAsync(dst, st.host, bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipEventRecord(st.used, b->stream));
  st.pending = true;
  return SCMF_OK;
}

// Workgroups for a grid-stride phase over `work` items: ~8 items per lane.
uint32_t stride_blocks(int64_t work) {
  int64_t nb = (work + BLOCK * 8 - 1) / (BLOCK * 8);
  return (uint32_t)std::max<int64_t>(1, std::min<int64_t>(nb, 1 << 16));
}

size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// What one membership needs: the workgroups of each sweep shape, with the
// longest chain among its members, and of the clear, gridding and omega
// phases; members ordered by decreasing beads, so the largest start first.
struct BatchNeed {
  std::vector<const SweepVariant*> v;
  std::vector<int> nmax;
  std::vector<std::vector<BlockRef>> sweep;
  std::vector<BlockRef> zero, grid, omega;
  bool tiled = false;  // a member grids in tiles (its own launches)
};

BatchNeed batch_need(scmf_batch* b) {
  BatchNeed nd;
  const int n = (int)b->m.size();
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(),
            [&](int a, int c) { return b->m[a]->n_beads > b->m[c]->n_beads; });
  for (int i : order) {
    scmf_sys* s = b->m[i];
    const SweepVariant* v = s->variant;
    if (b->forced_L) {
      const int32_t keep = s->opts.lanes_per_chain;
      s->opts.lanes_per_chain = b->forced_L;
      const SweepVariant* keepv = s->variant;
      const size_t keepl = s->lds_bytes;
      if (choose_sweep(s) == SCMF_OK) v = s->variant;
      s->variant = keepv;
      s->lds_bytes = keepl;
      s->opts.lanes_per_chain = keep;
    }
    size_t gi = 0;
    while (gi < nd.v.size() && nd.v[gi] != v) ++gi;
    if (gi == nd.v.size()) {
      nd.v.push_back(v);
      nd.nmax.push_back(0);
      nd.sweep.emplace_back();
    }
    nd.nmax[gi] = std::max(nd.nmax[gi], s->nmax);
    const int G = v->L == 0 ? BLOCK : 64 / v->L;
    const uint32_t nb = (uint32_t)((s->n_chains + G - 1) / G);
    for (uint32_t k = 0; k < nb; ++k) nd.sweep[gi].push_back({(uint32_t)i, k, nb, 0});
  }
  for (int i : order) {
    scmf_sys* s = b->m[i];
    const uint32_t nz = stride_blocks(s->count_words);
    for (uint32_t k = 0; k < nz; ++k) nd.zero.push_back({(uint32_t)i, k, nz, 0});
    if (s->tile_lds && s->sorted) {
      nd.tiled = true;
    } else {
      const uint32_t ng = stride_blocks(s->n_beads);
      for (uint32_t k = 0; k < ng; ++k) nd.grid.push_back({(uint32_t)i, k, ng, 0});
    }
    const uint32_t no = stride_blocks(s->n_cells);
    for (uint32_t k = 0; k < no; ++k) nd.omega.push_back({(uint32_t)i, k, no, 0});
  }
  return nd;
}

int batch_group_of(const scmf_batch* b, const SweepVariant* v) {
  for (size_t k = 0; k < b->groups.size(); ++k)
    if (b->groups[k].v == v) return (int)k;
  return -1;
}

// Whether the captured graph can run this membership from a new table: every
// sweep shape captured with room for its workgroups and chain length, room in
// every phase, no tiled launches either side, and at least a quarter of the
// launched workgroups in use (else a smaller graph pays for its capture).
bool batch_fits(const scmf_batch* b, const BatchNeed& nd, int n) {
  if (!b->exec || b->tiled || nd.tiled || n > b->member_cap) return false;
  uint64_t used = nd.zero.size() + nd.grid.size() + nd.omega.size(), launched = 0;
  for (size_t k = 0; k < nd.v.size(); ++k) {
    const int gi = batch_group_of(b, nd.v[k]);
    if (gi < 0) return false;
    const scmf_batch::Group& g = b->groups[gi];
    if (nd.nmax[k] > g.nmax || nd.sweep[k].size() > g.nblocks) return false;
    used += nd.sweep[k].size();
  }
  if (nd.zero.size() > b->n_zero || nd.grid.size() > b->n_grid || nd.omega.size() > b->n_omega)
    return false;
  for (const scmf_batch::Group& g : b->groups) launched += g.nblocks;
  launched += (uint64_t)b->n_zero + b->n_grid + b->n_omega;
  return launched <= 4096 || 4 * used >= launched;
}

// Fill the current generation's table for this membership (vacant entries
// past it) and upload it on the batch stream.
int batch_install(scmf_batch* b, const BatchNeed& nd) {
  const int n = (int)b->m.size();
  std::vector<uint8_t> blob(b->buf_bytes, 0);
  KParams* ps = (KParams*)blob.data();
  for (int i = 0; i < n; ++i) {
    ps[i] = b->m[i]->kp;
    ps[i].cells = nullptr;  // the batch grids from positions
    ps[i].sched = nullptr;  // and schedules inside the sweep
    ps[i].stop = b->stop[i];
  }
  KParams vacant{};
  vacant.stop = 0;  // stopped at sweep word 0, step word 0: every phase returns at once
  vacant.ctr = (uint32_t*)(b->d_buf + b->vac_ctr);
  for (int i = n; i <= b->member_cap; ++i) ps[i] = vacant;
  const BlockRef none{(uint32_t)b->member_cap, 0, 1, 0};
  auto put = [&](size_t off, unsigned cap, const std::vector<BlockRef>& t) {
    BlockRef* dst = (BlockRef*)(blob.data() + off);
    for (unsigned k = 0; k < cap; ++k) dst[k] = k < t.size() ? t[k] : none;
  };
  uint64_t used = nd.zero.size() + nd.grid.size() + nd.omega.size();
  static const std::vector<BlockRef> empty;
  for (const scmf_batch::Group& g : b->groups) {
    size_t k = 0;
    while (k < nd.v.size() && nd.v[k] != g.v) ++k;
    put(g.bt, g.nblocks, k < nd.v.size() ? nd.sweep[k] : empty);
    if (k < nd.v.size()) used += nd.sweep[k].size();
  }
  put(b->bt_zero, b->n_zero, nd.zero);
  put(b->bt_grid, b->n_grid, nd.grid);
  put(b->bt_omega, b->n_omega, nd.omega);
  b->used_blocks = used;
  return batch_upload(b, b->d_buf, blob.data(), blob.size());
}

// Workgroups a capacity holds: the need, with half again as much room when
// the membership is expected to change.
unsigned batch_room(size_t need, bool room) {
  return (unsigned)std::max<size_t>(1, room ? need + need / 2 + 64 : need);
}

// Install the current membership: into the captured graph's table when it
// fits, else as a new generation (room: capacities with headroom).
int batch_build(scmf_batch* b, bool room) {
  batch_reclaim(b, false);
  const BatchNeed nd = batch_need(b);
  const int n = (int)b->m.size();
  if (batch_fits(b, nd, n)) {
    int rc = batch_install(b, nd);
    if (!rc) b->uploads += 1;
    return rc;
  }
  int rc = batch_retire(b);
  if (rc) return rc;
  if (n == 0) return SCMF_OK;
  // the shape of the new generation
  b->member_cap = room ? n + n / 2 + 4 : n;
  b->tiled = nd.tiled;
  for (size_t k = 0; k < nd.v.size(); ++k) {
    scmf_batch::Group g{nd.v[k], nd.nmax[k], 0, 0, batch_room(nd.sweep[k].size(), room)};
    // rows rounded up to 16 with room, unless that no longer fits the LDS
    if (room && nd.v[k]->L != 0) {
      const int up = (nd.nmax[k] + 15) & ~15;
      if (up == nd.nmax[k] || nd.v[k]->bytes(up) <= 65536 ||
          raise_lds_limit(nd.v[k]->fn, nd.v[k]->bytes(up)) == hipSuccess)
        g.nmax = up;
      else
        (void)hipGetLastError();
    }
    g.lds = g.v->L == 0 ? 0 : g.v->bytes(g.nmax);
    if (g.lds > 65536 && raise_lds_limit(g.v->fn, g.lds) != hipSuccess) {
      (void)hipGetLastError();
      return fail(SCMF_E_ARG, "batch members' chains are too long for one LDS tile shape");
    }
    b->groups.push_back(g);
  }
  b->n_zero = batch_room(nd.zero.size(), room);
  b->n_grid = nd.grid.empty() && !room ? 0 : batch_room(nd.grid.size(), room);
  b->n_omega = batch_room(nd.omega.size(), room);
  // layout: parameter records, tables, the vacant record's sweep words
  size_t off = align16(sizeof(KParams) * (size_t)(b->member_cap + 1));
  for (scmf_batch::Group& g : b->groups) {
    g.bt = off;
    off = align16(off + sizeof(BlockRef) * g.nblocks);
  }
  b->bt_zero = off;
  off = align16(off + sizeof(BlockRef) * b->n_zero);
  b->bt_grid = off;
  off = align16(off + sizeof(BlockRef) * std::max(1u, b->n_grid));
  b->bt_omega = off;
  off = align16(off + sizeof(BlockRef) * b->n_omega);
  b->vac_ctr = off;
  b->buf_bytes = off + 16;
  HIP_TRY(hipMallocAsync((void**)&b->d_buf, b->buf_bytes, b->stream));
  if ((rc = batch_install(b, nd))) return rc;
  // capture one step on the capture stream
  const KParams* Ps = (const KParams*)b->d_buf;
  auto table = [&](size_t o) { return (const BlockRef*)(b->d_buf + o); };
  std::lock_guard<std::mutex> capture_lock(scmf_batch_tickets::capture_mutex());
  HIP_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
  for (const scmf_batch::Group& g : b->groups) {
    if (g.v->L == 0) {
      hipLaunchKernelGGL(k_sweep_global, dim3(g.nblocks), dim3(BLOCK), 0, b->cap, Ps, table(g.bt), 0,
                         (int64_t)0, INT64_MAX);
    } else {
      const BlockRef* bt = table(g.bt);
      int nmax = g.nmax, clear = 0;
      int64_t c0 = 0, c1 = INT64_MAX;  // every chain of each member
      void* args[] = {(void*)&Ps, (void*)&bt, (void*)&nmax, (void*)&clear, (void*)&c0, (void*)&c1};
      (void)hipLaunchKernel(g.v->fn, dim3(g.nblocks), dim3(64), args, g.lds, b->cap);
    }
  }
  const int nk = b->member_cap + 1;
  hipLaunchKernelGGL(k_advance, dim3((unsigned)((nk + 63) / 64)), dim3(64), 0, b->cap, Ps, nk);
  hipLaunchKernelGGL(k_zero_counts, dim3(b->n_zero), dim3(BLOCK), 0, b->cap, Ps, table(b->bt_zero));
  if (b->n_grid)
    hipLaunchKernelGGL(k_grid, dim3(b->n_grid), dim3(BLOCK), 0, b->cap, Ps, table(b->bt_grid), 0);
  int trc = SCMF_OK;
  for (int i = 0; i < n && trc == SCMF_OK; ++i)
    if (b->m[i]->tile_lds && b->m[i]->sorted) trc = launch_tiled_grid(b->m[i], b->cap, Ps, i, false, 0);
  hipLaunchKernelGGL(k_omega, dim3(b->n_omega), dim3(BLOCK), 0, b->cap, Ps, table(b->bt_omega), 1);
  hipError_t e = hipStreamEndCapture(b->cap, &b->graph);
  if (trc) return trc;
  if (e != hipSuccess) return fail(SCMF_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->generations += 1;
  return SCMF_OK;
}

// Members must be idle on their own streams and, if gridded in tiles, ordered.
int batch_prepare_members(scmf_sys* const* members, int32_t n) {
  for (int32_t i = 0; i < n; ++i) {
    scmf_sys* s = members[i];
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->tile_lds && !s->sorted && s->sort_every > 0) {
      int rc = sort_chains(s);
      if (rc) return rc;
      HIP_TRY(hipStreamSynchronize(s->stream));
    }
  }
  return SCMF_OK;
}

int batch_check_members(scmf_sys* const* members, int32_t n) {
  if (n < 0 || (n > 0 && !members)) return fail(SCMF_E_ARG, "batch needs members");
  for (int32_t i = 0; i < n; ++i) {
    if (!members[i]) return fail(SCMF_E_ARG, "null batch member");
    for (int32_t j = 0; j < i; ++j)
      if (members[j] == members[i]) return fail(SCMF_E_ARG, "a system appears twice in the batch");
  }
  return SCMF_OK;
}

uint32_t stop_of(int64_t v) {
  return v < 0 || v >= (int64_t)NO_STOP ? NO_STOP : (uint32_t)v;
}

// The launches of n steps on the batch stream (host bookkeeping included).
int batch_launch(scmf_batch* b, int32_t n_steps) {
  batch_reclaim(b, false);
  for (int32_t k = 0; k < n_steps; ++k) {
    // spatial reordering of tile-gridded members, between graph replays
    bool reordered = false;
    for (size_t i = 0; i < b->m.size(); ++i) {
      scmf_sys* s = b->m[i];
      if (s->sweeps >= b->stop[i]) continue;
      if (s->tile_lds && s->sort_every > 0 && s->sweeps % (uint64_t)s->sort_every == 0 &&
          s->sweeps > 0) {
        HIP_TRY(hipStreamSynchronize(b->stream));
        int rc = sort_chains(s);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(s->stream));
        reordered = true;
      }
    }
    if (reordered) {
      int rc = batch_build(b, false);
      if (rc) return rc;
    }
    if (!b->exec) break;  // no members
    HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    for (size_t i = 0; i < b->m.size(); ++i) {
      scmf_sys* s = b->m[i];
      s->cells_fresh = false;
      if (s->sweeps < b->stop[i]) {
        s->sweeps += 1;
        s->moves += (uint64_t)s->n_beads;
      }
    }
    b->steps += 1;
  }
  return SCMF_OK;
}

}  // namespace

// Steps of a batch behind a ticket, for the job server: events around them on
// the batch stream, read without waiting.
namespace scmf_batch_tickets {

int enqueue(scmf_batch* b, int32_t n_steps, uint64_t* ticket) {
  scmf_batch::Ticket t{b->next_ticket, nullptr, nullptr};
  HIP_TRY(hipEventCreate(&t.a));
  HIP_TRY(hipEventCreate(&t.b));
  HIP_TRY(hipEventRecord(t.a, b->stream));
  int rc = batch_launch(b, n_steps);
  if (rc) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
    return rc;
  }
  HIP_TRY(hipEventRecord(t.b, b->stream));
  b->next_ticket += 1;
  b->tickets.push_back(t);
  *ticket = t.id;
  return SCMF_OK;
}

// *done = 1 once the steps of `ticket` have run (wait: block until then);
// *t_end = when they ended, in seconds after the batch's origin event;
// *busy = their device time.  A finished ticket is forgotten after this.
int poll(scmf_batch* b, uint64_t ticket, bool wait, int* done, double* t_end, double* busy) {
  *done = 0;
  for (auto it = b->tickets.begin(); it != b->tickets.end(); ++it) {
    if (it->id != ticket) continue;
    const hipError_t q = wait ? hipEventSynchronize(it->b) : hipEventQuery(it->b);
    if (q == hipErrorNotReady) return SCMF_OK;
    HIP_TRY(q);
    float ms_end = 0, ms_busy = 0;
    HIP_TRY(hipEventElapsedTime(&ms_end, b->origin, it->b));
    HIP_TRY(hipEventElapsedTime(&ms_busy, it->a, it->b));
    *t_end = ms_end * 1e-3;
    *busy = ms_busy * 1e-3;
    (void)hipEventDestroy(it->a);
    (void)hipEventDestroy(it->b);
    b->tickets.erase(it);
    *done = 1;
    return SCMF_OK;
  }
  return fail(SCMF_E_ARG, "unknown ticket");
}

std::chrono::steady_clock::time_point origin(scmf_batch* b) { return b->origin_host; }

std::mutex& capture_mutex() {
  static std::mutex mu;
  return mu;
}

}  // namespace scmf_batch_tickets

extern "C" {

int scmf_batch_create(scmf_batch** out, scmf_sys* const* members, int32_t n_members) {
  if (!out) return fail(SCMF_E_ARG, "out is null");
  *out = nullptr;
  if (n_members <= 0) return fail(SCMF_E_ARG, "batch needs members");
  int rc = batch_check_members(members, n_members);
  if (rc) return rc;
  scmf_batch* b = new scmf_batch();
  b->m.assign(members, members + n_members);
  b->stop.assign(n_members, NO_STOP);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->origin) != hipSuccess || hipEventRecord(b->origin, b->stream) != hipSuccess ||
      hipEventSynchronize(b->origin) != hipSuccess) {
    (void)hipGetLastError();
    rc = fail(SCMF_E_HIP, "batch stream set-up failed");
  }
  b->origin_host = std::chrono::steady_clock::now();
  if (!rc) rc = batch_prepare_members(members, n_members);
  if (!rc) rc = batch_build(b, false);
  if (rc) {
    scmf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return SCMF_OK;
}

int scmf_batch_update(scmf_batch* b, scmf_sys* const* members, int32_t n_members,
                      const int64_t* stop_sweeps) {
  if (!b) return fail(SCMF_E_ARG, "null batch");
  int rc = batch_check_members(members, n_members);
  if (rc) return rc;
  // members that are new to the batch must be idle on their own streams
  std::vector<scmf_sys*> fresh;
  for (int32_t i = 0; i < n_members; ++i)
    if (std::find(b->m.begin(), b->m.end(), members[i]) == b->m.end()) fresh.push_back(members[i]);
  if ((rc = batch_prepare_members(fresh.data(), (int32_t)fresh.size()))) return rc;
  b->m.assign(members, members + n_members);
  b->stop.assign(n_members, NO_STOP);
  if (stop_sweeps)
    for (int32_t i = 0; i < n_members; ++i) b->stop[i] = stop_of(stop_sweeps[i]);
  return batch_build(b, true);
}

void scmf_batch_destroy(scmf_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  (void)batch_retire(b);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_reclaim(b, true);
  for (auto& t : b->tickets) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto& st : b->stage) {
    if (st.host) (void)hipHostFree(st.host);
    if (st.used) (void)hipEventDestroy(st.used);
  }
  if (b->origin) (void)hipEventDestroy(b->origin);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int scmf_batch_step(scmf_batch* b, int32_t n_steps) {
  if (!b) return fail(SCMF_E_ARG, "null batch");
  if (n_steps < 0) return fail(SCMF_E_ARG, "n_steps must be non-negative");
  return batch_launch(b, n_steps);
}

int scmf_batch_sync(scmf_batch* b) {
  if (!b) return fail(SCMF_E_ARG, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  batch_reclaim(b, false);
  return SCMF_OK;
}

int scmf_batch_set_lanes(scmf_batch* b, int32_t lanes_per_chain) {
  if (!b) return fail(SCMF_E_ARG, "null batch");
  b->forced_L = lanes_per_chain;
  return batch_build(b, false);
}

int scmf_batch_stats(scmf_batch* b, scmf_batch_stats_t* out) {
  if (!b || !out) return fail(SCMF_E_ARG, "null argument");
  memset(out, 0, sizeof *out);
  out->steps = b->steps;
  out->captures = b->generations;
  out->table_uploads = b->uploads;
  out->members = (int32_t)b->m.size();
  out->member_capacity = b->member_cap;
  uint64_t launched = (uint64_t)b->n_zero + b->n_grid + b->n_omega;
  for (const scmf_batch::Group& g : b->groups) launched += g.nblocks;
  out->blocks_launched = b->exec ? launched : 0;
  out->blocks_used = b->exec ? b->used_blocks : 0;
  return SCMF_OK;
}

}  // extern "C"
